"""Loads fastload_amd/libfastload.so (built in-tree by fastload_amd/build.py) and declares the C
ABI of include/fastload.h and include/fastload_gen.h for ctypes. There is no fallback: if the
library is missing or fails to load, importing this module raises."""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libfastload.so"

i32p = C.POINTER(C.c_int)
f64p = C.POINTER(C.c_double)
vp = C.c_void_p


class SlideMatrix(C.Structure):
    """slide_matrix (src/format.h:6-20; include/fastload.h)."""
    _fields_ = [
        ("sortval", f64p), ("sortrowidx", i32p), ("tile_ptr", i32p), ("tile_len", i32p),
        ("tile_colidx", i32p), ("tile_tall", i32p), ("tile_format", i32p), ("countrow", i32p),
        ("segmentoffset", i32p), ("sortrowindex", i32p), ("tilenum", C.c_int), ("segsum", C.c_int),
    ]


class MatrixInfo(C.Structure):
    _fields_ = [(k, C.c_int) for k in ("m", "n", "nnz", "width", "levels", "tilenum", "segsum", "max_col_nnz")]


# name -> (restype, argtypes)
SIGNATURES = {
    "fl_last_error": (C.c_char_p, []),
    "fl_mmio_allinone": (C.c_int, [i32p, i32p, i32p, i32p, C.POINTER(i32p), C.POINTER(i32p), C.POINTER(f64p), C.c_char_p]),
    "fl_free": (None, [vp]),
    "fl_csr_to_csc": (C.c_int, [f64p, C.c_int, C.c_int, C.c_int, vp, vp, vp, vp, vp, vp, vp]),
    "fl_ColSort": (C.c_int, [f64p, C.c_int, C.c_int, C.c_int, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
    "fl_formatTransform": (C.c_int, [f64p, f64p, C.POINTER(SlideMatrix), vp, vp, vp, C.c_int, C.c_int, C.c_int, i32p]),
    "fl_FastLoad_spmv": (C.c_int, [C.c_char_p, C.POINTER(SlideMatrix), C.c_int, C.c_int, C.c_int, vp, vp, vp, f64p]),
    "fl_Tile_destroy": (None, [C.POINTER(SlideMatrix)]),
    "fl_slide_save": (C.c_int, [C.c_char_p, C.POINTER(SlideMatrix), C.c_int, C.c_int, C.c_int, C.c_int, vp]),
    "fl_slide_load": (C.c_int, [C.c_char_p, C.POINTER(SlideMatrix), i32p, i32p, i32p, i32p, C.POINTER(i32p)]),
    "fl_matrix_save": (C.c_int, [vp, C.c_char_p, vp]),
    "fl_matrix_load": (C.c_int, [C.POINTER(vp), C.c_char_p, vp]),
    "fl_FastLoad_cpu": (C.c_int, [C.POINTER(SlideMatrix), C.c_int, C.c_int, C.c_int, vp, vp, vp, C.c_int, i32p]),
    "fl_FastLoad_cpu_f32": (C.c_int, [C.POINTER(SlideMatrix), vp, C.c_int, C.c_int, C.c_int, vp, vp, C.c_int]),
    "fl_device_count": (C.c_int, [i32p]),
    "fl_malloc": (C.c_int, [C.POINTER(vp), C.c_size_t]),
    "fl_free_device": (C.c_int, [vp]),
    "fl_memcpy_h2d": (C.c_int, [vp, vp, C.c_size_t]),
    "fl_memcpy_d2h": (C.c_int, [vp, vp, C.c_size_t]),
    "fl_memcpy_d2d": (C.c_int, [vp, vp, C.c_size_t]),
    "fl_memset_device": (C.c_int, [vp, C.c_int, C.c_size_t]),
    "fl_synchronize": (C.c_int, []),
    "fl_stream_create": (C.c_int, [C.POINTER(vp)]),
    "fl_stream_destroy": (C.c_int, [vp]),
    "fl_stream_synchronize": (C.c_int, [vp]),
    "fl_event_create": (C.c_int, [C.POINTER(vp)]),
    "fl_event_destroy": (C.c_int, [vp]),
    "fl_event_record": (C.c_int, [vp, vp]),
    "fl_event_elapsed": (C.c_int, [C.POINTER(C.c_float), vp, vp]),
    "fl_matrix_create": (C.c_int, [C.POINTER(vp), C.c_int, C.c_int, C.c_int, vp, vp, vp, C.c_int, vp]),
    "fl_matrix_destroy": (C.c_int, [vp]),
    "fl_matrix_get_info": (C.c_int, [vp, C.POINTER(MatrixInfo)]),
    "fl_matrix_export": (C.c_int, [vp] + [vp] * 14),
    "fl_permute_x_f64": (C.c_int, [vp, vp, vp, vp]),
    "fl_permute_x_f32": (C.c_int, [vp, vp, vp, vp]),
    "fl_matrix_prepare_values": (C.c_int, [vp, C.c_int, vp]),
    "fl_matrix_pack_info": (C.c_int, [vp, C.c_int, i32p, i32p, i32p]),
    "fl_spmv": (C.c_int, [vp, C.c_int, C.c_int, vp, vp, vp, vp]),
    "fl_spmv_assign": (C.c_int, [vp, C.c_int, C.c_int, vp, vp, vp, vp]),
    "fl_set_stream_policy": (C.c_int, [C.c_int]),
    "fl_spmv_assign_noalloc": (C.c_int, [vp, C.c_int, C.c_int, vp, vp, vp]),
    "fl_matrix_prepare_schedule": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, vp]),
    "fl_matrix_prepare_hyb": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, vp]),
    "fl_matrix_hyb_info": (C.c_int, [vp, i32p, C.POINTER(C.c_longlong)]),
    "fl_matrix_prepare_rows": (C.c_int, [vp, vp]),
    "fl_spmm": (C.c_int, [vp, C.c_int, C.c_int, vp, vp, C.c_int, vp]),
    "fl_batch_assemble": (C.c_int, [C.c_int, vp, vp, vp, vp, vp, i32p, i32p, C.POINTER(i32p), C.POINTER(i32p),
                                    C.POINTER(f64p), vp, vp]),
    "fl_graph_begin_capture": (C.c_int, [vp]),
    "fl_graph_end_capture": (C.c_int, [vp, C.POINTER(vp)]),
    "fl_graph_launch": (C.c_int, [vp, vp]),
    "fl_graph_destroy": (C.c_int, [vp]),
    "fl_zero_device": (C.c_int, [vp, C.c_size_t, vp]),
    "fl_cg_create": (C.c_int, [C.POINTER(vp), vp, C.c_int, C.c_int]),
    "fl_cg_solve": (C.c_int, [vp, vp, vp, C.c_int, C.c_double, C.c_int, i32p, C.POINTER(C.c_double)]),
    "fl_cg_destroy": (C.c_int, [vp]),
    "fl_features": (C.c_int, [C.c_int, C.c_int, C.c_int, vp, vp, vp, vp]),
    "fl_mlp_create": (C.c_int, [C.POINTER(vp), C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64]),
    "fl_mlp_destroy": (C.c_int, [vp]),
    "fl_mlp_param_count": (C.c_int, [vp]),
    "fl_mlp_get_params": (C.c_int, [vp, vp]),
    "fl_mlp_set_params": (C.c_int, [vp, vp]),
    "fl_mlp_forward": (C.c_int, [vp, vp, C.c_int, vp, vp]),
    "fl_mlp_predict": (C.c_int, [vp, vp, C.c_int, vp, vp, vp]),
    "fl_mlp_gradients": (C.c_int, [vp, vp, vp, C.c_int, vp, C.POINTER(C.c_float), vp]),
    "fl_mlp_train_step": (C.c_int, [vp, vp, vp, C.c_int, C.c_float, C.POINTER(C.c_float), vp]),
    "fl_schedule_info": (C.c_int, [C.c_int, i32p, i32p, i32p, i32p]),
    "fl_select_schedule": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, vp, vp, i32p, C.POINTER(C.c_float), vp]),
    "fl_gen_matrix": (C.c_int, [C.c_int, C.c_uint64, C.c_int, C.c_int, vp, C.c_int, C.POINTER(i32p), C.POINTER(i32p),
                                C.POINTER(f64p), i32p]),
    "fl_gen_vector": (C.c_int, [C.c_int, C.c_uint64, C.c_int, vp]),
}


def _load():
    if not LIB_PATH.exists():
        raise ImportError(f"fastload_amd: native library {LIB_PATH} is missing; build it with "
                          f"`python fastload_amd/build.py` (hipcc, gfx950)")
    lib = C.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | C.RTLD_GLOBAL)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


class FastLoadError(RuntimeError):
    """A library call failed (message from fl_last_error), or a wrapper refused its arguments
    before calling into the library (rc -8, FL_ERR_INVALID, with the wrapper's own message)."""

    def __init__(self, rc: int, where: str, from_library: bool = True):
        msg = lib.fl_last_error().decode(errors="replace") if from_library else ""
        super().__init__(f"{where} failed with status {rc}: {msg}" if from_library else f"{where} (status {rc})")
        self.rc = rc


def check(rc: int, where: str) -> int:
    if rc < 0:
        raise FastLoadError(rc, where)
    return rc
