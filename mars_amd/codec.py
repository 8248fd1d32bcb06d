"""Host-side encode path above the search provider (include/mars_codec.h).

Mirrors crates/mars-codec/src/encode.rs:
  * EncodeOptions            76-101
  * encode_image_with_search 1054-1126 (+ mars_format::write, mars_format.rs:232-297)
  * color::wrap_gray_stream  (color.rs:382-384)
  * metrics::ycbcr (metrics.rs:377-417), color::downsample_box (color.rs:60-78),
    the MARC container (color.rs:778-790) and the colour encode
    (color.rs:400-489) with a per-plane provider

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


class CColorStats(C.Structure):
    _fields_ = [("y_bytes", C.c_uint64), ("cb_bytes", C.c_uint64), ("cr_bytes", C.c_uint64),
                ("y", CEncodeStats), ("cb", CEncodeStats), ("cr", CEncodeStats)]


COLOR_GRAY, COLOR_RGB_444, COLOR_RGB_420 = 0, 1, 2

# (user, row, col, size, shift, max, out[max], n_out, evals) -> status
SEARCH_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                        C.c_uint32, C.POINTER(N.CCandidate), C.POINTER(C.c_uint32),
                        C.POINTER(C.c_uint64))


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
        L.mars_hip_encode_plane_method.argtypes = [
            C.c_void_p, P(C.c_uint8), C.c_uint32, C.c_uint32, P(CEncodeParams),
            P(CEncodeOptions), C.c_uint32, P(P(C.c_uint8)), P(C.c_size_t), P(CEncodeStats)]
        L.mars_hip_encode_plane_method.restype = C.c_int
        L.mars_ycbcr.argtypes = [P(C.c_uint8), C.c_size_t, P(C.c_uint8), P(C.c_uint8),
                                 P(C.c_uint8)]
        L.mars_ycbcr.restype = C.c_int
        L.mars_downsample_box.argtypes = [P(C.c_uint8), C.c_uint32, C.c_uint32, P(C.c_uint8)]
        L.mars_downsample_box.restype = C.c_int
        L.mars_write_container.argtypes = [C.c_uint8, P(P(C.c_uint8)), P(C.c_size_t),
                                           C.c_uint32, P(P(C.c_uint8)), P(C.c_size_t)]
        L.mars_write_container.restype = C.c_int
        L.mars_hip_encode_rgb.argtypes = [
            C.c_void_p, P(C.c_uint8), C.c_uint32, C.c_uint32, C.c_uint32, P(CEncodeParams),
            P(CEncodeParams), P(CEncodeOptions), C.c_uint32, P(P(C.c_uint8)), P(C.c_size_t),
            P(CColorStats)]
        L.mars_hip_encode_rgb.restype = C.c_int
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


def _u8p(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_uint8))


def ycbcr(rgb: np.ndarray) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """metrics::ycbcr of an h x w x 3 u8 image: (Y, Cb, Cr) planes."""
    L = _lib()
    img = np.ascontiguousarray(rgb, dtype=np.uint8)
    if img.ndim != 3 or img.shape[2] != 3:
        raise ValueError("rgb must be h x w x 3")
    h, w, _ = img.shape
    y, cb, cr = (np.zeros((h, w), np.uint8) for _ in range(3))
    if L.mars_ycbcr(_u8p(img), h * w, _u8p(y), _u8p(cb), _u8p(cr)) != N.OK:
        raise RuntimeError("mars_ycbcr failed")
    return y, cb, cr


def downsample_box(plane: np.ndarray) -> np.ndarray:
    """color::downsample_box: the 4:2:0 chroma plane, ceil(h/2) x ceil(w/2)."""
    L = _lib()
    p = np.ascontiguousarray(plane, dtype=np.uint8)
    h, w = p.shape
    out = np.zeros(((h + 1) // 2, (w + 1) // 2), np.uint8)
    if L.mars_downsample_box(_u8p(p), w, h, _u8p(out)) != N.OK:
        raise RuntimeError("mars_downsample_box failed")
    return out


def write_container(mode: int, streams) -> bytes:
    """The MARC container (color.rs:778-790) of the given plane streams."""
    L = _lib()
    bufs = [(C.c_uint8 * max(len(b), 1)).from_buffer_copy(b or b"\0") for b in streams]
    ptrs = (C.POINTER(C.c_uint8) * max(len(bufs), 1))(
        *[C.cast(b, C.POINTER(C.c_uint8)) for b in bufs])
    lens = (C.c_size_t * max(len(bufs), 1))(*[len(b) for b in streams])
    buf = C.POINTER(C.c_uint8)()
    n = C.c_size_t()
    if L.mars_write_container(mode, ptrs, lens, len(bufs), C.byref(buf), C.byref(n)) != N.OK:
        raise RuntimeError("mars_write_container failed")
    return _take(L, buf, n)


def hip_encode_plane_method(ctx: HipContext, image: np.ndarray, params: EncodeParams,
                            options: EncodeOptions, method: int) -> Tuple[bytes, EncodeStats]:
    """encode_image_with_search with IndexedSearchProvider::build(.., method) on the GPU."""
    L = _lib()
    img = np.ascontiguousarray(image, dtype=np.uint8)
    h, w = img.shape
    buf = C.POINTER(C.c_uint8)()
    n = C.c_size_t()
    st = CEncodeStats()
    rc = L.mars_hip_encode_plane_method(ctx._ctx, _u8p(img), w, h, C.byref(_params(params)),
                                        C.byref(_options(options)), method, C.byref(buf),
                                        C.byref(n), C.byref(st))
    ctx._check(rc, "mars_hip_encode_plane_method")
    return _take(L, buf, n), _stats(st)


def hip_encode_rgb(ctx: HipContext, rgb: np.ndarray, y_params: EncodeParams,
                   chroma_params: EncodeParams, options: EncodeOptions, method: int,
                   subsampling: int = COLOR_RGB_420):
    """Colour encode (color.rs:400-489) with `method` as every plane's provider.

    Returns (MARC bytes, {"y": EncodeStats, "cb": ..., "cr": ...}, per-plane byte counts).
    """
    L = _lib()
    img = np.ascontiguousarray(rgb, dtype=np.uint8)
    h, w, _ = img.shape
    buf = C.POINTER(C.c_uint8)()
    n = C.c_size_t()
    st = CColorStats()
    rc = L.mars_hip_encode_rgb(ctx._ctx, _u8p(img), w, h, subsampling,
                               C.byref(_params(y_params)), C.byref(_params(chroma_params)),
                               C.byref(_options(options)), method, C.byref(buf), C.byref(n),
                               C.byref(st))
    ctx._check(rc, "mars_hip_encode_rgb")
    stats = {"y": _stats(st.y), "cb": _stats(st.cb), "cr": _stats(st.cr)}
    return _take(L, buf, n), stats, (st.y_bytes, st.cb_bytes, st.cr_bytes)
