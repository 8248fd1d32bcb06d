"""Host-side encode path above the search provider (include/mars_codec.h).

Mirrors crates/mars-codec/src/encode.rs:
  * EncodeOptions            76-101
  * encode_image_with_search 1054-1126 (+ mars_format::write, mars_format.rs:232-297)
  * color::wrap_gray_stream  (color.rs:382-384)

``encode_plane(image, params, options, search, warmup)`` takes the search
provider as a C callback (``mars_search_fn``); ``hip_encode_plane`` uses the
GPU search table of a ``HipContext`` for both the RD warm-up and the walk.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional, Tuple

import numpy as np

from . import _native as N
from .provider import EncodeParams, HipContext


class CEncodeParams(C.Structure):
    _fields_ = [("min_size", C.c_uint32), ("max_size", C.c_uint32), ("shift", C.c_uint32),
                ("bits_alfa", C.c_uint32), ("bits_beta", C.c_uint32),
                ("zero_threshold", C.c_uint32), ("max_alfa", C.c_double), ("t_rms", C.c_double),
                ("has_lambda", C.c_int32), ("_pad", C.c_int32), ("lambda_", C.c_double)]


class CEncodeOptions(C.Structure):
    _fields_ = [("allowed_modes", C.c_uint8 * 4), ("adaptive_density", C.c_uint8),
                ("lambda_adaptive_qstep", C.c_uint8), ("_pad", C.c_uint8 * 2),
                ("fixed_qstep", C.c_double), ("rd_candidates", C.c_uint32),
                ("_pad2", C.c_uint32)]


class CEncodeStats(C.Structure):
    _fields_ = [("search_evals", C.c_uint64), ("warmup_evals", C.c_uint64),
                ("leaf_modes", C.c_uint64 * 4), ("split_decisions", C.c_uint64),
                ("leaf_decisions", C.c_uint64), ("n_leaves", C.c_uint64)]


SEARCH_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                        C.POINTER(N.CCandidate), C.POINTER(C.c_uint32), C.POINTER(C.c_uint64))


@dataclass(frozen=True)
class EncodeOptions:
    """encode.rs:76-101 (no lambda regions; one RD candidate)."""
    allowed_modes: Tuple[bool, bool, bool, bool] = (True, True, True, True)
    adaptive_density: bool = False
    lambda_adaptive_qstep: bool = False   # ResidualQuantisation::LambdaAdaptive
    fixed_qstep: float = 8.0              # ResidualQuantisation::Fixed (LEGACY = 8)
    rd_candidates: int = 1


@dataclass
class EncodeStats:
    search_evals: int
    warmup_evals: int
    leaf_modes: Tuple[int, int, int, int]
    split_decisions: int
    leaf_decisions: int
    n_leaves: int


def _lib():
    L = N.lib()
    if not getattr(L, "_codec_bound", False):
        P = C.POINTER
        L.mars_encode_plane.argtypes = [P(C.c_uint8), C.c_uint32, C.c_uint32, P(CEncodeParams),
                                        P(CEncodeOptions), C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_void_p, P(P(C.c_uint8)), P(C.c_size_t),
                                        P(CEncodeStats)]
        L.mars_encode_plane.restype = C.c_int
        L.mars_hip_encode_plane.argtypes = [C.c_void_p, P(C.c_uint8), C.c_uint32, C.c_uint32,
                                            P(CEncodeParams), P(CEncodeOptions),
                                            P(P(C.c_uint8)), P(C.c_size_t), P(CEncodeStats)]
        L.mars_hip_encode_plane.restype = C.c_int
        L.mars_hip_encode_prepared.argtypes = [C.c_void_p, C.c_uint32, P(CEncodeParams),
                                               P(CEncodeOptions), P(P(C.c_uint8)),
                                               P(C.c_size_t), P(CEncodeStats)]
        L.mars_hip_encode_prepared.restype = C.c_int
        L.mars_wrap_gray.argtypes = [P(C.c_uint8), C.c_size_t, P(P(C.c_uint8)), P(C.c_size_t)]
        L.mars_wrap_gray.restype = C.c_int
        L.mars_free.argtypes = [C.c_void_p]
        L.mars_free.restype = None
        L._codec_bound = True
    return L


def _params(p: EncodeParams) -> CEncodeParams:
    return CEncodeParams(p.min_size, p.max_size, p.shift, p.bits_alfa, p.bits_beta,
                         p.zero_threshold, p.max_alfa, p.t_rms, int(p.lambda_ is not None), 0,
                         float(p.lambda_ or 0.0))


def _options(o: EncodeOptions) -> CEncodeOptions:
    c = CEncodeOptions()
    for i, m in enumerate(o.allowed_modes):
        c.allowed_modes[i] = int(bool(m))
    c.adaptive_density = int(o.adaptive_density)
    c.lambda_adaptive_qstep = int(o.lambda_adaptive_qstep)
    c.fixed_qstep = o.fixed_qstep
    c.rd_candidates = o.rd_candidates
    return c


def _take(L, buf, n) -> bytes:
    data = C.string_at(buf, n.value)
    L.mars_free(buf)
    return data


def _stats(s: CEncodeStats) -> EncodeStats:
    return EncodeStats(s.search_evals, s.warmup_evals, tuple(s.leaf_modes), s.split_decisions,
                       s.leaf_decisions, s.n_leaves)


def encode_plane(image: np.ndarray, params: EncodeParams, options: EncodeOptions,
                 search: C.c_void_p, search_user: C.c_void_p,
                 warmup: Optional[C.c_void_p] = None,
                 warmup_user: Optional[C.c_void_p] = None) -> Tuple[bytes, EncodeStats]:
    """Encode one grey plane to `.mars` bytes with the given search callbacks."""
    L = _lib()
    img = np.ascontiguousarray(image, dtype=np.uint8)
    h, w = img.shape
    buf = C.POINTER(C.c_uint8)()
    n = C.c_size_t()
    st = CEncodeStats()
    rc = L.mars_encode_plane(img.ctypes.data_as(C.POINTER(C.c_uint8)), w, h,
                             C.byref(_params(params)), C.byref(_options(options)), search,
                             search_user, warmup if warmup is not None else search,
                             warmup_user if warmup is not None else search_user, C.byref(buf),
                             C.byref(n), C.byref(st))
    if rc != N.OK:
        raise RuntimeError(f"mars_encode_plane failed with status {rc}")
    return _take(L, buf, n), _stats(st)


def hip_encode_plane(ctx: HipContext, image: np.ndarray, params: EncodeParams,
                     options: EncodeOptions) -> Tuple[bytes, EncodeStats]:
    """The same encode with the context's GPU search table as the provider."""
    L = _lib()
    img = np.ascontiguousarray(image, dtype=np.uint8)
    h, w = img.shape
    buf = C.POINTER(C.c_uint8)()
    n = C.c_size_t()
    st = CEncodeStats()
    rc = L.mars_hip_encode_plane(ctx._ctx, img.ctypes.data_as(C.POINTER(C.c_uint8)), w, h,
                                 C.byref(_params(params)), C.byref(_options(options)),
                                 C.byref(buf), C.byref(n), C.byref(st))
    ctx._check(rc, "mars_hip_encode_plane")
    return _take(L, buf, n), _stats(st)


def hip_encode_prepared(ctx: HipContext, plane: int, params: EncodeParams,
                        options: EncodeOptions) -> Tuple[bytes, EncodeStats]:
    """Encode plane `plane` of the batch prepared in ctx from its search table.

    Runs without the GIL (ctypes releases it), so several planes can be
    encoded from a thread pool at once.
    """
    L = _lib()
    buf = C.POINTER(C.c_uint8)()
    n = C.c_size_t()
    st = CEncodeStats()
    rc = L.mars_hip_encode_prepared(ctx._ctx, plane, C.byref(_params(params)),
                                    C.byref(_options(options)), C.byref(buf), C.byref(n),
                                    C.byref(st))
    ctx._check(rc, "mars_hip_encode_prepared")
    return _take(L, buf, n), _stats(st)


def wrap_gray(plane_bytes: bytes) -> bytes:
    """color::wrap_gray_stream: the MARC container of a grey `.mars` plane."""
    L = _lib()
    src = (C.c_uint8 * max(len(plane_bytes), 1)).from_buffer_copy(plane_bytes or b"\0")
    buf = C.POINTER(C.c_uint8)()
    n = C.c_size_t()
    rc = L.mars_wrap_gray(src, len(plane_bytes), C.byref(buf), C.byref(n))
    if rc != N.OK:
        raise RuntimeError(f"mars_wrap_gray failed with status {rc}")
    return _take(L, buf, n)
