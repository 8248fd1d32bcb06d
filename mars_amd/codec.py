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


# Typed errors of the readers (mars_format.rs:141-181 MarsFormatError, ifs.rs:130-156).
FMT_E_BAD_MAGIC, FMT_E_INVALID_QSTEP, FMT_E_UNSUPPORTED_VERSION, FMT_E_TRUNCATED = -101, -102, -103, -104
FMT_E_DEGENERATE_HEADER, FMT_E_SECTION_OUT_OF_BOUNDS, FMT_E_MISSING_TREE = -105, -106, -107
FMT_E_INVALID_TREE_PAYLOAD, FMT_E_DOMAIN_OUT_OF_BOUNDS, FMT_E_TRAILING_BITS = -108, -109, -110
FMT_E_NONZERO_PADDING, FMT_E_UNSUPPORTED = -111, -112


class CHeader(C.Structure):
    """mars_format.rs:59-64 Header (geometry + residual step bits)."""
    _fields_ = [(n, C.c_uint32) for n in ("min_size", "max_size", "shift", "bits_alfa",
                                          "bits_beta", "int_max_alfa", "width", "height",
                                          "residual_qstep_bits", "_pad")]


class CLeaf(C.Structure):
    """ifs.rs:100-117 Leaf (residual levels in a separate array)."""
    _fields_ = [(n, C.c_uint32) for n in ("row", "col", "size", "mode", "qalfa", "qbeta",
                                          "isometry", "dom_row", "dom_col")] + \
               [("qgx", C.c_int32), ("qgy", C.c_int32), ("residual_off", C.c_uint32),
                ("residual_len", C.c_uint32)]


class FormatError(ValueError):
    """A reader rejected its input; .code is the MARS_FMT_E_* status."""

    def __init__(self, code: int):
        super().__init__(f"stream rejected with status {code}")
        self.code = code


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
        L.mars_read.argtypes = [P(C.c_uint8), C.c_size_t, P(CHeader), P(P(CLeaf)),
                                P(C.c_size_t), P(P(C.c_int32)), P(C.c_size_t)]
        L.mars_read.restype = C.c_int
        L.mars_write.argtypes = [P(CHeader), P(CLeaf), C.c_size_t, P(C.c_int32), C.c_size_t,
                                 P(P(C.c_uint8)), P(C.c_size_t)]
        L.mars_write.restype = C.c_int
        L.mars_ifs_parse.argtypes = [P(C.c_uint8), C.c_size_t, P(CHeader), P(P(CLeaf)),
                                     P(C.c_size_t)]
        L.mars_ifs_parse.restype = C.c_int
        L.mars_decode_iterative.argtypes = [P(CHeader), P(CLeaf), C.c_size_t, P(C.c_int32),
                                            C.c_size_t, C.c_uint32, P(C.c_uint8)]
        L.mars_decode_iterative.restype = C.c_int
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


# ---- reading streams back ---------------------------------------------------------

def _leaf_tuple(c: CLeaf, res):
    """(row, col, size, mode, qalfa, qbeta, isometry, dom_row, dom_col, qgx, qgy, residual)."""
    r = tuple(res[c.residual_off:c.residual_off + c.residual_len]) if c.residual_len else ()
    return (c.row, c.col, c.size, c.mode, c.qalfa, c.qbeta, c.isometry, c.dom_row, c.dom_col,
            c.qgx, c.qgy, r)


def _header_dict(h: CHeader) -> dict:
    return {n: getattr(h, n) for n, _ in CHeader._fields_ if n != "_pad"}


def read_mars(data: bytes):
    """mars_format::read: (header dict, leaves as tuples -- see _leaf_tuple) or FormatError."""
    L = _lib()
    buf = (C.c_uint8 * max(len(data), 1)).from_buffer_copy(data or b"\0")
    hdr = CHeader()
    lp = C.POINTER(CLeaf)()
    rp = C.POINTER(C.c_int32)()
    n, nr = C.c_size_t(), C.c_size_t()
    rc = L.mars_read(buf, len(data), C.byref(hdr), C.byref(lp), C.byref(n), C.byref(rp),
                     C.byref(nr))
    if rc != N.OK:
        raise FormatError(rc)
    try:
        res = [rp[i] for i in range(nr.value)]
        leaves = [_leaf_tuple(lp[i], res) for i in range(n.value)]
    finally:
        L.mars_free(C.cast(lp, C.c_void_p))
        L.mars_free(C.cast(rp, C.c_void_p))
    return _header_dict(hdr), leaves


def _c_leaves(leaves):
    arr = (CLeaf * max(len(leaves), 1))()
    res = []
    for i, t in enumerate(leaves):
        row, col, size, mode, qa, qb, iso, dr, dc, gx, gy, r = t
        arr[i] = CLeaf(row, col, size, mode, qa, qb, iso, dr, dc, gx, gy, len(res), len(r))
        res.extend(r)
    rarr = (C.c_int32 * max(len(res), 1))(*res)
    return arr, rarr, len(res)


def write_mars(header: dict, leaves) -> bytes:
    """mars_format::write of a header dict and leaf tuples (as read_mars returns them)."""
    L = _lib()
    h = CHeader(**header)
    arr, rarr, nres = _c_leaves(leaves)
    out = C.POINTER(C.c_uint8)()
    n = C.c_size_t()
    rc = L.mars_write(C.byref(h), arr, len(leaves), rarr, nres, C.byref(out), C.byref(n))
    if rc != N.OK:
        raise FormatError(rc)
    return _take(L, out, n)


def ifs_parse(data: bytes):
    """ifs::parse of a legacy Mars 1 `.ifs` bitstream: (header dict, leaf tuples)."""
    L = _lib()
    buf = (C.c_uint8 * max(len(data), 1)).from_buffer_copy(data or b"\0")
    hdr = CHeader()
    lp = C.POINTER(CLeaf)()
    n = C.c_size_t()
    rc = L.mars_ifs_parse(buf, len(data), C.byref(hdr), C.byref(lp), C.byref(n))
    if rc != N.OK:
        raise FormatError(rc)
    try:
        leaves = [_leaf_tuple(lp[i], []) for i in range(n.value)]
    finally:
        L.mars_free(C.cast(lp, C.c_void_p))
    return _header_dict(hdr), leaves


def decode_iterative(header: dict, leaves, iterations: int) -> np.ndarray:
    """ifs::decode_iterative: `iterations` fixed-point iterations from flat 128."""
    L = _lib()
    h = CHeader(**header)
    arr, rarr, nres = _c_leaves(leaves)
    out = np.zeros((header["height"], header["width"]), np.uint8)
    rc = L.mars_decode_iterative(C.byref(h), arr, len(leaves), rarr, nres, iterations,
                                 out.ctypes.data_as(C.POINTER(C.c_uint8)))
    if rc != N.OK:
        raise ValueError(f"mars_decode_iterative failed with status {rc}")
    return out
