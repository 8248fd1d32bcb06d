"""ctypes binding of the CPU oracle (oracle/liboracle.so) -- test infrastructure.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
always as the checker.
"""

from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
LIB = ROOT / "oracle" / "liboracle.so"


class Moments(C.Structure):
    _fields_ = [("s0", C.c_int64), ("s1_x4", C.c_int64), ("s2_x16", C.c_int64),
                ("t0", C.c_int64), ("t1_x4", C.c_int64), ("t2", C.c_int64)]


class OrCandidate(C.Structure):
    _fields_ = [("dom_row", C.c_uint32), ("dom_col", C.c_uint32), ("isometry", C.c_uint32),
                ("qalfa", C.c_uint32), ("qbeta", C.c_uint32), ("_pad", C.c_uint32),
                ("rms", C.c_double), ("m", Moments)]


class FisherEntry(C.Structure):
    _fields_ = [("dom_row", C.c_uint32), ("dom_col", C.c_uint32), ("dom_iso", C.c_uint32)]


NB = 72  # Fisher buckets: 3 shape classes x 24 variance orders


class FisherIndex(C.Structure):
    _fields_ = [("size", C.c_uint32), ("shift", C.c_uint32), ("P", C.c_uint32),
                ("start", C.c_uint32 * NB), ("count", C.c_uint32 * NB),
                ("own_count", C.c_uint32 * NB), ("entries", C.POINTER(FisherEntry))]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB.exists():
        subprocess.run(["make", "-C", str(LIB.parent)], check=True, capture_output=True)
    L = C.CDLL(str(LIB))
    P, u32, u64, i64, dbl = C.POINTER, C.c_uint32, C.c_uint64, C.c_int64, C.c_double
    L.or_contract.argtypes = [P(C.c_uint8), u32, u32, P(C.c_int32)]
    L.or_contract.restype = None
    L.or_fit_f64.argtypes = [P(Moments), dbl, u32, u32, P(u32), P(u32), P(dbl)]
    L.or_fit_f64.restype = None
    L.or_fit_f32.argtypes = [P(Moments), C.c_float, u32, u32, P(u32), P(u32), P(C.c_float)]
    L.or_fit_f32.restype = None
    L.or_best_beta.argtypes = [i64, i64, u32]
    L.or_best_beta.restype = u32
    L.or_quantise_f64.argtypes = [dbl, u32]
    L.or_quantise_f64.restype = u32
    L.or_iso_map.argtypes = [u32, u32, u32, u32, P(u32), P(u32)]
    L.or_iso_map.restype = None
    L.or_domain_sums.argtypes = [P(C.c_int32), C.c_size_t, C.c_size_t, C.c_size_t, C.c_size_t,
                                 P(i64), P(i64)]
    L.or_domain_sums.restype = None
    L.or_permute_range.argtypes = [P(C.c_uint8), u32, u32, P(C.c_uint8)]
    L.or_permute_range.restype = None
    L.or_cross_term_permuted.argtypes = [P(C.c_int32), C.c_size_t, C.c_size_t, C.c_size_t,
                                         C.c_size_t, P(C.c_uint8)]
    L.or_cross_term_permuted.restype = i64
    L.or_search_top_k.argtypes = [P(C.c_uint8), u32, u32, P(C.c_int32), u32, u32, u32, u32, dbl,
                                  u32, u32, u32, P(OrCandidate), P(u32), P(u64)]
    L.or_search_top_k.restype = None
    L.or_search_table.argtypes = [P(C.c_uint8), u32, u32, u32, u32, dbl, u32, u32, P(u32),
                                  P(u32), u32, C.c_int, P(OrCandidate), P(C.c_uint8)]
    L.or_search_table.restype = u64
    L.or_search_table_k.argtypes = [P(C.c_uint8), u32, u32, u32, u32, dbl, u32, u32, P(u32),
                                    P(u32), u32, C.c_int, u32, P(OrCandidate), P(C.c_uint8)]
    L.or_search_table_k.restype = u64
    L.or_provider_new.argtypes = [P(C.c_uint8), u32, u32, dbl, u32, u32]
    L.or_provider_new.restype = C.c_void_p
    L.or_provider_free.argtypes = [C.c_void_p]
    L.or_provider_free.restype = None
    D = P(dbl)
    L.or_fisher_tables.argtypes = [P(C.c_int32), P(C.c_uint8)]
    L.or_fisher_tables.restype = None
    L.or_fisher_newclass.argtypes = [D, u32, P(u32), P(u32)]
    L.or_fisher_newclass.restype = None
    L.or_fisher_variance_class.argtypes = [D, u32]
    L.or_fisher_variance_class.restype = u32
    L.or_fisher_flips.argtypes = [D, u32, u32, D]
    L.or_fisher_flips.restype = None
    L.or_fisher_classify.argtypes = [D, u32, P(u32), P(u32), P(u32)]
    L.or_fisher_classify.restype = None
    L.or_fisher_index_build.argtypes = [P(C.c_int32), u32, u32, u32, u32]
    L.or_fisher_index_build.restype = P(FisherIndex)
    L.or_fisher_index_free.argtypes = [P(FisherIndex)]
    L.or_fisher_index_free.restype = None
    L.or_fisher_table.argtypes = [P(C.c_uint8), u32, u32, u32, u32, dbl, u32, u32, P(u32),
                                  P(u32), u32, C.c_int, P(OrCandidate), P(C.c_uint8), P(u64)]
    L.or_fisher_table.restype = u64
    L.or_fisher_provider_new.argtypes = [P(C.c_uint8), u32, u32, dbl, u32, u32]
    L.or_fisher_provider_new.restype = C.c_void_p
    L.or_fisher_provider_free.argtypes = [C.c_void_p]
    L.or_fisher_provider_free.restype = None
    L.or_ycbcr.argtypes = [P(C.c_uint8), C.c_size_t, P(C.c_uint8), P(C.c_uint8), P(C.c_uint8)]
    L.or_ycbcr.restype = None
    L.or_downsample_box.argtypes = [P(C.c_uint8), u32, u32, P(C.c_uint8)]
    L.or_downsample_box.restype = None
    _lib = L
    return L


def fit_f64(m, max_alfa=1.0, bits_alfa=4, bits_beta=7):
    L = lib()
    mm = Moments(*[int(x) for x in m])
    qa, qb, rms = C.c_uint32(), C.c_uint32(), C.c_double()
    L.or_fit_f64(C.byref(mm), max_alfa, bits_alfa, bits_beta, C.byref(qa), C.byref(qb),
                 C.byref(rms))
    return qa.value, qb.value, rms.value


def fit_f32(m, max_alfa=1.0, bits_alfa=4, bits_beta=7):
    L = lib()
    mm = Moments(*[int(x) for x in m])
    qa, qb, rms = C.c_uint32(), C.c_uint32(), C.c_float()
    L.or_fit_f32(C.byref(mm), max_alfa, bits_alfa, bits_beta, C.byref(qa), C.byref(qb),
                 C.byref(rms))
    return qa.value, qb.value, rms.value


def permute_range(rng: np.ndarray, k: int) -> np.ndarray:
    rng = np.ascontiguousarray(rng, np.uint8)
    s = rng.shape[0]
    out = np.zeros_like(rng)
    lib().or_permute_range(rng.ctypes.data_as(C.POINTER(C.c_uint8)), k, s,
                           out.ctypes.data_as(C.POINTER(C.c_uint8)))
    return out


def integer_moments(img: np.ndarray, row: int, col: int, size: int, dom_row: int, dom_col: int,
                    k: int):
    """RawMoments of one candidate through the oracle's own kernels."""
    img = np.ascontiguousarray(img, np.uint8)
    d = contract(img)
    L = lib()
    s1, s2 = C.c_int64(), C.c_int64()
    dr, dc = dom_row // 2, dom_col // 2
    dp = d.ctypes.data_as(C.POINTER(C.c_int32))
    L.or_domain_sums(dp, d.shape[1], dr, dc, size, C.byref(s1), C.byref(s2))
    rng = img[row:row + size, col:col + size]
    rk = permute_range(rng, k)
    t1 = L.or_cross_term_permuted(dp, d.shape[1], dr, dc, size,
                                  rk.ctypes.data_as(C.POINTER(C.c_uint8)))
    r = rng.astype(np.int64)
    return (size * size, s1.value, s2.value, int(r.sum()), int(t1), int((r * r).sum()))


def contract(img: np.ndarray) -> np.ndarray:
    img = np.ascontiguousarray(img, np.uint8)
    h, w = img.shape
    out = np.zeros((h // 2, w // 2), np.int32)
    lib().or_contract(img.ctypes.data_as(C.POINTER(C.c_uint8)), w, h,
                      out.ctypes.data_as(C.POINTER(C.c_int32)))
    return out


def search_top_k(img, row, col, size, shift, k=1, max_alfa=1.0, bits_alfa=4, bits_beta=7):
    img = np.ascontiguousarray(img, np.uint8)
    h, w = img.shape
    d = contract(img)
    out = (OrCandidate * max(k, 1))()
    n, ev = C.c_uint32(), C.c_uint64()
    lib().or_search_top_k(img.ctypes.data_as(C.POINTER(C.c_uint8)), w, h,
                          d.ctypes.data_as(C.POINTER(C.c_int32)), row, col, size, shift,
                          max_alfa, bits_alfa, bits_beta, k, out, C.byref(n), C.byref(ev))
    return [out[i] for i in range(n.value)], ev.value


def search_table(img, size, shift, rows, cols, threads=8, max_alfa=1.0, bits_alfa=4,
                 bits_beta=7):
    img = np.ascontiguousarray(img, np.uint8)
    h, w = img.shape
    n = len(rows)
    r = (C.c_uint32 * max(n, 1))(*rows)
    c = (C.c_uint32 * max(n, 1))(*cols)
    out = (OrCandidate * max(n, 1))()
    found = (C.c_uint8 * max(n, 1))()
    ev = lib().or_search_table(img.ctypes.data_as(C.POINTER(C.c_uint8)), w, h, size, shift,
                               max_alfa, bits_alfa, bits_beta, r, c, n, threads, out, found)
    return [out[i] if found[i] else None for i in range(n)], ev


def search_table_k(img, size, shift, rows, cols, k, threads=8, max_alfa=1.0, bits_alfa=4,
                   bits_beta=7):
    """Each range's best k (search_top_k): a list of candidate lists, and evals."""
    img = np.ascontiguousarray(img, np.uint8)
    h, w = img.shape
    n = len(rows)
    r = (C.c_uint32 * max(n, 1))(*rows)
    c = (C.c_uint32 * max(n, 1))(*cols)
    out = (OrCandidate * max(n * k, 1))()
    found = (C.c_uint8 * max(n, 1))()
    ev = lib().or_search_table_k(img.ctypes.data_as(C.POINTER(C.c_uint8)), w, h, size, shift,
                                 max_alfa, bits_alfa, bits_beta, r, c, n, threads, k, out, found)
    return [[out[i * k + j] for j in range(found[i])] for i in range(n)], ev


def nodes(w, h, min_size, max_size):
    """The searchable nodes of a quadtree walk (encode.rs:1330-1362, 1865-1919)."""
    out = {}

    def visit(row, col, size):
        if row >= h or col >= w:
            return
        forced = size > max_size or row + size > h or col + size > w
        if forced:
            half = size // 2
            for r, c in ((row, col), (row + half, col), (row, col + half),
                         (row + half, col + half)):
                visit(r, c, half)
            return
        if size == 1:
            return
        out.setdefault(size, []).append((row, col))
        if size > min_size:
            half = size // 2
            for r, c in ((row, col), (row + half, col), (row, col + half),
                         (row + half, col + half)):
                visit(r, c, half)

    vs = 1
    while vs < max(w, h):
        vs *= 2
    visit(0, 0, vs)
    return out


class Provider:
    """The oracle's exhaustive search as a mars_search_fn callback + user pointer."""

    def __init__(self, img, max_alfa=1.0, bits_alfa=4, bits_beta=7):
        L = lib()
        img = np.ascontiguousarray(img, np.uint8)
        h, w = img.shape
        self._L = L
        self.user = C.c_void_p(L.or_provider_new(img.ctypes.data_as(C.POINTER(C.c_uint8)), w, h,
                                                 max_alfa, bits_alfa, bits_beta))
        self.fn = C.cast(L.or_search_cb, C.c_void_p)

    def close(self):
        if self.user:
            self._L.or_provider_free(self.user)
            self.user = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- Fisher classified search and colour plumbing ----------------------------

def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _u8p(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_uint8))


def fisher_tables():
    o = (C.c_int32 * 144)()
    m = (C.c_uint8 * 64)()
    lib().or_fisher_tables(o, m)
    return np.array(o, np.int32).reshape(24, 6), np.array(m, np.uint8).reshape(8, 8)


def fisher_newclass(block: np.ndarray):
    b = np.ascontiguousarray(block, np.float64)
    iso, clas = C.c_uint32(), C.c_uint32()
    lib().or_fisher_newclass(_dp(b), b.shape[0], C.byref(iso), C.byref(clas))
    return iso.value, clas.value


def fisher_variance_class(block: np.ndarray) -> int:
    b = np.ascontiguousarray(block, np.float64)
    return lib().or_fisher_variance_class(_dp(b), b.shape[0])


def fisher_flips(block: np.ndarray, iso: int) -> np.ndarray:
    b = np.ascontiguousarray(block, np.float64)
    out = np.zeros_like(b)
    lib().or_fisher_flips(_dp(b), b.shape[0], iso, _dp(out))
    return out


def fisher_classify(block: np.ndarray):
    """(isometry, class, variance class) of one block (fisher.rs:35-38)."""
    b = np.ascontiguousarray(block, np.float64)
    iso, clas, vc = C.c_uint32(), C.c_uint32(), C.c_uint32()
    lib().or_fisher_classify(_dp(b), b.shape[0], C.byref(iso), C.byref(clas), C.byref(vc))
    return iso.value, clas.value, vc.value


def fisher_index(img: np.ndarray, size: int, shift: int):
    """Buckets of one (size, stride): (list of 72 entry lists, own counts)."""
    img = np.ascontiguousarray(img, np.uint8)
    h, w = img.shape
    d = contract(img)
    L = lib()
    ix = L.or_fisher_index_build(d.ctypes.data_as(C.POINTER(C.c_int32)), w, h, size, shift)
    try:
        x = ix.contents
        buckets = [[(x.entries[i].dom_row, x.entries[i].dom_col, x.entries[i].dom_iso)
                    for i in range(x.start[b], x.start[b] + x.count[b])] for b in range(NB)]
        own = list(x.own_count)
    finally:
        L.or_fisher_index_free(ix)
    return buckets, own


def fisher_table(img, size, shift, rows, cols, threads=8, max_alfa=1.0, bits_alfa=4,
                 bits_beta=7):
    """Fisher winners (None where no candidate) and per-range evals."""
    img = np.ascontiguousarray(img, np.uint8)
    h, w = img.shape
    n = len(rows)
    r = (C.c_uint32 * max(n, 1))(*rows)
    c = (C.c_uint32 * max(n, 1))(*cols)
    out = (OrCandidate * max(n, 1))()
    found = (C.c_uint8 * max(n, 1))()
    each = (C.c_uint64 * max(n, 1))()
    lib().or_fisher_table(_u8p(img), w, h, size, shift, max_alfa, bits_alfa, bits_beta, r, c, n,
                          threads, out, found, each)
    return [out[i] if found[i] else None for i in range(n)], [each[i] for i in range(n)]


class FisherProvider:
    """The oracle's Fisher search as a mars_search_fn callback + user pointer."""

    def __init__(self, img, max_alfa=1.0, bits_alfa=4, bits_beta=7):
        L = lib()
        img = np.ascontiguousarray(img, np.uint8)
        h, w = img.shape
        self._L = L
        self.user = C.c_void_p(L.or_fisher_provider_new(_u8p(img), w, h, max_alfa, bits_alfa,
                                                        bits_beta))
        self.fn = C.cast(L.or_fisher_cb, C.c_void_p)

    def close(self):
        if self.user:
            self._L.or_fisher_provider_free(self.user)
            self.user = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ycbcr(rgb: np.ndarray):
    rgb = np.ascontiguousarray(rgb, np.uint8)
    h, w, _ = rgb.shape
    y, cb, cr = (np.zeros((h, w), np.uint8) for _ in range(3))
    lib().or_ycbcr(_u8p(rgb), h * w, _u8p(y), _u8p(cb), _u8p(cr))
    return y, cb, cr


def downsample_box(plane: np.ndarray) -> np.ndarray:
    plane = np.ascontiguousarray(plane, np.uint8)
    h, w = plane.shape
    out = np.zeros(((h + 1) // 2, (w + 1) // 2), np.uint8)
    lib().or_downsample_box(_u8p(plane), w, h, _u8p(out))
    return out
