"""MI355X-native FastLoad: CSC SpMV with columns sorted by non-zero count and wave-wide tiles,
hand-written HIP kernels for gfx950 behind a C ABI (include/fastload.h).

Importing this package loads fastload_amd/libfastload.so; if it is missing the import fails.
"""
from ._lib import FastLoadError, lib  # noqa: F401  (raises ImportError when the .so is missing)
from .api import (  # noqa: F401
    GEN_BANDED, GEN_BLOCKDIAG, GEN_STENCIL27, GEN_SYMGEOM, GEN_UNIFORM, GEN_ZIPF, GEN_ZIPF_DENSE, KERNEL_DROPIN,
    KERNEL_FAST, KERNEL_HYB, KERNEL_PB, KERNEL_WIN, KERNEL_WINR, VAL_BF16, VAL_BF16X2, VAL_F32, VAL_F64, VAL_PACKED, VEC_MOD10, VEC_NORMAL, VEC_UNIFORM, ColSort,
    CG, KERNEL_ROW, FastLoad_cpu, FastLoad_cpu_f32, FastLoad_spmv, slide_load, slide_save, FastLoadMatrix, Graph, Tiles, batch_assemble, csr_to_csc, formatTransform, gen_matrix, gen_vector,
    mmio_allinone, zero_device, STREAM_AUTO, STREAM_NEVER, STREAM_ALWAYS, set_stream_policy)
from .device import DeviceArray, Event, Stream, device_count, synchronize  # noqa: F401

__all__ = [n for n in dir() if not n.startswith("_")]
