"""ctypes binding of ``libmars_hip.so`` (the C ABI in include/mars_hip.h).

Loading fails loudly when the library is missing: there is no CPU fallback.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libmars_hip.so"

OK = 0
E_ARG, E_DEVICE, E_NOT_PREPARED, E_NOT_IN_TABLE, E_UNSUPPORTED = -1, -2, -3, -4, -5
KERNEL_MFMA_I8, KERNEL_FP32 = 0, 1
METHOD_EXHAUSTIVE, METHOD_FISHER = 0, 1

# Every symbol include/mars_hip.h declares.
EXPORTS = (
    "mars_hip_open", "mars_hip_close", "mars_hip_prepare", "mars_hip_prepare_batch",
    "mars_hip_search", "mars_hip_search_plane", "mars_hip_last_error", "mars_hip_set_kernel",
    "mars_hip_run", "mars_hip_autotune", "mars_hip_table_count", "mars_hip_table_plane",
    "mars_hip_table_info", "mars_hip_table_read", "mars_hip_table_scores", "mars_hip_fit_f64",
    "mars_hip_copy_probe", "mars_hip_search_method", "mars_hip_table_read_method",
    "mars_hip_set_ranked", "mars_hip_table_read_ranked", "mars_hip_set_walk_tables",
    "mars_hip_search_grid_noalloc",
)


class Params(C.Structure):
    _fields_ = [("min_size", C.c_uint32), ("max_size", C.c_uint32), ("shift", C.c_uint32),
                ("bits_alfa", C.c_uint32), ("bits_beta", C.c_uint32), ("method", C.c_uint32),
                ("max_alfa", C.c_double)]


class CCandidate(C.Structure):
    _fields_ = [("dom_row", C.c_uint32), ("dom_col", C.c_uint32), ("isometry", C.c_uint32),
                ("qalfa", C.c_uint32), ("qbeta", C.c_uint32), ("_pad", C.c_uint32),
                ("rms", C.c_double), ("s0", C.c_int64), ("s1_x4", C.c_int64),
                ("s2_x16", C.c_int64), ("t0", C.c_int64), ("t1_x4", C.c_int64),
                ("t2", C.c_int64)]


class Timing(C.Structure):
    _fields_ = [("ms_step", C.c_double), ("ms_pool", C.c_double), ("ms_search", C.c_double),
                ("ms_by_size", C.c_double * 6), ("evals", C.c_uint64),
                ("survivors", C.c_uint64), ("survivors_by_size", C.c_uint64 * 6),
                ("exact_fits_by_size", C.c_uint64 * 6), ("ms_fisher", C.c_double),
                ("fisher_evals", C.c_uint64), ("ms_pool_build", C.c_double)]


class Plane(C.Structure):
    _fields_ = [("px", C.POINTER(C.c_uint8)), ("w", C.c_uint32), ("h", C.c_uint32)]


class Tuning(C.Structure):
    _fields_ = [("target_blocks", C.c_uint32 * 6), ("min_tiles", C.c_uint32 * 6),
                ("ms", C.c_double * 6)]


_lib = None


def lib() -> C.CDLL:
    """Load libmars_hip.so (raises if it is missing)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
    L = C.CDLL(str(LIB_PATH))
    P, u32, i32, u64 = C.POINTER, C.c_uint32, C.c_int, C.c_uint64
    L.mars_hip_open.argtypes = [i32, P(C.c_void_p)]
    L.mars_hip_close.argtypes = [C.c_void_p]
    L.mars_hip_close.restype = None
    L.mars_hip_prepare.argtypes = [C.c_void_p, P(C.c_uint8), u32, u32, P(Params), P(u32), u32]
    L.mars_hip_prepare_batch.argtypes = [C.c_void_p, P(Plane), u32, P(Params), P(u32), u32]
    L.mars_hip_search.argtypes = [C.c_void_p, u32, u32, u32, u32, u32, P(CCandidate), P(u32),
                                  P(u64)]
    L.mars_hip_search_plane.argtypes = [C.c_void_p, u32, u32, u32, u32, u32, u32, P(CCandidate),
                                        P(u32), P(u64)]
    L.mars_hip_autotune.argtypes = [C.c_void_p, u32, P(Tuning)]
    L.mars_hip_table_plane.argtypes = [C.c_void_p, u32, P(u32)]
    L.mars_hip_last_error.argtypes = [C.c_void_p]
    L.mars_hip_last_error.restype = C.c_char_p
    L.mars_hip_set_kernel.argtypes = [C.c_void_p, i32]
    L.mars_hip_run.argtypes = [C.c_void_p, i32, i32, P(Timing)]
    L.mars_hip_table_count.argtypes = [C.c_void_p, P(u32)]
    L.mars_hip_table_info.argtypes = [C.c_void_p, u32, P(u32), P(u32), P(u32), P(u32)]
    L.mars_hip_table_read.argtypes = [C.c_void_p, u32, P(u32), P(u32), P(CCandidate),
                                      P(C.c_uint8)]
    L.mars_hip_table_scores.argtypes = [C.c_void_p, u32, P(C.c_double)]
    L.mars_hip_fit_f64.argtypes = [C.c_void_p, P(C.c_int64), u32, C.c_double, u32, u32, P(u32),
                                   P(u32), P(C.c_double)]
    L.mars_hip_copy_probe.argtypes = [C.c_void_p, C.c_size_t, P(C.c_double)]
    L.mars_hip_search_method.argtypes = [C.c_void_p, u32, u32, u32, u32, u32, u32, u32,
                                         P(CCandidate), P(u32), P(u64)]
    L.mars_hip_table_read_method.argtypes = [C.c_void_p, u32, u32, P(CCandidate), P(C.c_uint8),
                                             P(u64)]
    L.mars_hip_set_ranked.argtypes = [C.c_void_p, u32]
    L.mars_hip_set_walk_tables.argtypes = [C.c_void_p, C.c_double]
    L.mars_hip_table_read_ranked.argtypes = [C.c_void_p, u32, P(u32), P(CCandidate),
                                             P(C.c_uint8)]
    L.mars_hip_search_grid_noalloc.argtypes = [C.c_void_p, C.c_void_p, u32, u32, P(Params),
                                               C.c_void_p, C.c_void_p, C.c_void_p]
    for name in EXPORTS:
        if name not in ("mars_hip_close", "mars_hip_last_error"):
            getattr(L, name).restype = C.c_int
    _lib = L
    return L
