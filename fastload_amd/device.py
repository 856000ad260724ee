"""Device buffers, streams and events over the library's own HIP runtime (no torch)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib


def device_count() -> int:
    c = C.c_int(0)
    check(lib.fl_device_count(C.byref(c)), "fl_device_count")
    return c.value


class DeviceArray:
    """A 1-D device allocation with a numpy dtype; freed when garbage-collected."""

    def __init__(self, count: int, dtype):
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        self.nbytes = self.dtype.itemsize * self.count
        p = C.c_void_p()
        check(lib.fl_malloc(C.byref(p), max(self.nbytes, 1)), "fl_malloc")
        self.ptr = p.value

    @classmethod
    def from_host(cls, a) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        if a.size:
            check(lib.fl_memcpy_h2d(d.ptr, a.ctypes.data, d.nbytes), "fl_memcpy_h2d")
        return d

    @classmethod
    def zeros(cls, count: int, dtype) -> "DeviceArray":
        d = cls(count, dtype)
        d.zero()
        return d

    def zero(self):
        if self.nbytes:
            check(lib.fl_memset_device(self.ptr, 0, self.nbytes), "fl_memset_device")

    def to_host(self) -> np.ndarray:
        out = np.empty(self.count, self.dtype)
        if self.nbytes:
            check(lib.fl_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes), "fl_memcpy_d2h")
        return out

    def view(self, offset: int, count: int) -> "DeviceArray":
        """Elements [offset, offset + count) of this allocation, without a copy. The view does
        not own memory: keep the parent alive while the view is used."""
        if offset < 0 or count < 0 or offset + count > self.count:
            raise ValueError(f"view [{offset}, {offset + count}) outside {self.count} elements")
        v = object.__new__(DeviceArray)
        v.dtype, v.count, v.nbytes = self.dtype, int(count), self.dtype.itemsize * int(count)
        v.ptr = self.ptr + self.dtype.itemsize * int(offset)
        v._parent = self
        return v

    def copy_from(self, other: "DeviceArray"):
        assert other.nbytes == self.nbytes
        if self.nbytes:
            check(lib.fl_memcpy_d2d(self.ptr, other.ptr, self.nbytes), "fl_memcpy_d2d")

    def free(self):
        if getattr(self, "_parent", None) is not None:  # a view: the parent owns the memory
            self.ptr = None
            return
        if getattr(self, "ptr", None):
            lib.fl_free_device(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize():
    check(lib.fl_synchronize(), "fl_synchronize")


class Stream:
    def __init__(self):
        p = C.c_void_p()
        check(lib.fl_stream_create(C.byref(p)), "fl_stream_create")
        self.ptr = p.value

    def synchronize(self):
        check(lib.fl_stream_synchronize(self.ptr), "fl_stream_synchronize")

    def __del__(self):
        try:
            if self.ptr:
                lib.fl_stream_destroy(self.ptr)
        except Exception:
            pass


class Event:
    def __init__(self):
        p = C.c_void_p()
        check(lib.fl_event_create(C.byref(p)), "fl_event_create")
        self.ptr = p.value

    def record(self, stream=None):
        check(lib.fl_event_record(self.ptr, stream.ptr if stream is not None else None), "fl_event_record")

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        check(lib.fl_event_elapsed(C.byref(ms), self.ptr, end.ptr), "fl_event_elapsed")
        return ms.value

    def __del__(self):
        try:
            if self.ptr:
                lib.fl_event_destroy(self.ptr)
        except Exception:
            pass
