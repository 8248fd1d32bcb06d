"""Host-side mirror of the reference search-provider interface over the C ABI.

Reference: crates/mars-codec/src/encode.rs
  * EncodeParams          27-60
  * RawMoments, Candidate 224-244
  * SearchRequest         530-540
  * SearchOutcome         542-546
  * SearchProvider        557-579 (search / search_candidates)
  * encode_image_with_search's factory, called once per plane, 1054-1076

``HipSearchProvider(image, params)`` is that factory's product for the HIP
backend: construction runs ``mars_hip_prepare`` (pool + whole-plane search
table on the GPU), ``search`` is the thread-safe lookup.  Like the reference
interface, ``search`` never raises for a legal request; an illegal one (a
range the walk could not request) raises ``ValueError`` the way the
reference's misuse panics (crates/mars-search/src/lib.rs:520-527).
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import _native as N


@dataclass(frozen=True)
class EncodeParams:
    """encode.rs:27-60 (only the fields the search and walks read)."""
    min_size: int = 4
    max_size: int = 16
    shift: int = 4
    bits_alfa: int = 4
    bits_beta: int = 7
    max_alfa: float = 1.0
    t_rms: float = 8.0
    zero_threshold: int = 0
    lambda_: Optional[float] = None


@dataclass(frozen=True)
class RawMoments:
    """encode.rs:224-232; s1/s2/t1 are 4x/16x/4x their value (exact integers)."""
    s0: int
    s1_x4: int
    s2_x16: int
    t0: int
    t1_x4: int
    t2: int


@dataclass(frozen=True)
class Candidate:
    """encode.rs:234-244."""
    dom_row: int
    dom_col: int
    isometry: int
    qalfa: int
    qbeta: int
    rms: float
    moments: RawMoments


@dataclass(frozen=True)
class SearchRequest:
    """encode.rs:530-540 (the plane is the provider's own copy)."""
    row: int
    col: int
    size: int
    shift: int


@dataclass
class SearchOutcome:
    """encode.rs:542-546."""
    candidate: Optional[Candidate]
    evals: int


@dataclass
class RankedCandidates:
    """encode.rs:548-555."""
    candidates: list = field(default_factory=list)
    evals: int = 0


def _cand(c: N.CCandidate) -> Candidate:
    return Candidate(c.dom_row, c.dom_col, c.isometry, c.qalfa, c.qbeta, c.rms,
                     RawMoments(c.s0, c.s1_x4, c.s2_x16, c.t0, c.t1_x4, c.t2))


class HipContext:
    """Owns one ``mars_hip_ctx`` (one device stream, one prepared plane)."""

    def __init__(self, device: int = 0, kernel: int = N.KERNEL_MFMA_I8):
        self._L = N.lib()
        self._ctx = C.c_void_p()
        rc = self._L.mars_hip_open(device, C.byref(self._ctx))
        if rc != N.OK:
            raise RuntimeError(f"mars_hip_open({device}) failed with status {rc}")
        self.set_kernel(kernel)

    def _check(self, rc: int, what: str) -> None:
        if rc != N.OK:
            msg = self._L.mars_hip_last_error(self._ctx).decode(errors="replace")
            raise RuntimeError(f"{what} failed with status {rc}: {msg}")

    def set_kernel(self, kernel: int) -> None:
        self._check(self._L.mars_hip_set_kernel(self._ctx, kernel), "mars_hip_set_kernel")

    def set_ranked(self, k: int) -> None:
        """Candidates per range the next prepare keeps (search_top_k's k, 1..32)."""
        self._check(self._L.mars_hip_set_ranked(self._ctx, k), "mars_hip_set_ranked")

    def set_walk_tables(self, t_rms: float) -> None:
        """t_rms > 0: the next Fisher prepare computes exhaustive winners only for
        the nodes a threshold walk at t_rms visits (the RD warm-up's); 0: full."""
        self._check(self._L.mars_hip_set_walk_tables(self._ctx, t_rms), "mars_hip_set_walk_tables")

    def prepare(self, image: np.ndarray, params: EncodeParams, shifts: Sequence[int],
                method: int = N.METHOD_EXHAUSTIVE) -> None:
        img = np.ascontiguousarray(image, dtype=np.uint8)
        if img.ndim != 2:
            raise ValueError("image must be a 2-D u8 plane")
        h, w = img.shape
        p = N.Params(params.min_size, params.max_size, params.shift, params.bits_alfa,
                     params.bits_beta, method, params.max_alfa)
        sh = (C.c_uint32 * len(shifts))(*shifts)
        self._check(self._L.mars_hip_prepare(
            self._ctx, img.ctypes.data_as(C.POINTER(C.c_uint8)), w, h, C.byref(p), sh,
            len(shifts)), "mars_hip_prepare")

    def prepare_batch(self, images: Sequence[np.ndarray], params: EncodeParams,
                      shifts: Sequence[int], method: int = N.METHOD_EXHAUSTIVE) -> None:
        """One pool and search table per plane, searched by one batched step."""
        imgs = [np.ascontiguousarray(im, dtype=np.uint8) for im in images]
        if not imgs or any(im.ndim != 2 for im in imgs):
            raise ValueError("images must be a non-empty list of 2-D u8 planes")
        planes = (N.Plane * len(imgs))()
        for i, im in enumerate(imgs):
            planes[i].px = im.ctypes.data_as(C.POINTER(C.c_uint8))
            planes[i].h, planes[i].w = im.shape
        p = N.Params(params.min_size, params.max_size, params.shift, params.bits_alfa,
                     params.bits_beta, method, params.max_alfa)
        sh = (C.c_uint32 * len(shifts))(*shifts)
        self._check(self._L.mars_hip_prepare_batch(self._ctx, planes, len(imgs), C.byref(p), sh,
                                                   len(shifts)), "mars_hip_prepare_batch")

    def autotune(self, reps: int = 3) -> N.Tuning:
        t = N.Tuning()
        self._check(self._L.mars_hip_autotune(self._ctx, reps, C.byref(t)), "mars_hip_autotune")
        return t

    def run(self, iters: int = 1, use_graph: bool = True) -> N.Timing:
        t = N.Timing()
        self._check(self._L.mars_hip_run(self._ctx, iters, int(use_graph), C.byref(t)),
                    "mars_hip_run")
        return t

    def search_raw(self, row: int, col: int, size: int, shift: int, plane: int = 0):
        out = N.CCandidate()
        n = C.c_uint32()
        ev = C.c_uint64()
        rc = self._L.mars_hip_search_plane(self._ctx, plane, row, col, size, shift, 1,
                                           C.byref(out), C.byref(n), C.byref(ev))
        return rc, (out if n.value else None), ev.value

    def search_ranked(self, row: int, col: int, size: int, shift: int, k: int, plane: int = 0):
        """(status, up to k candidates ascending, evals): search_candidates."""
        out = (N.CCandidate * k)()
        n = C.c_uint32()
        ev = C.c_uint64()
        rc = self._L.mars_hip_search_plane(self._ctx, plane, row, col, size, shift, k, out,
                                           C.byref(n), C.byref(ev))
        return rc, [out[i] for i in range(n.value)], ev.value

    def ranked_tables(self, plane: int = 0):
        """{(size, shift): (rows, cols, [list of up to k candidates per range])}."""
        res = {}
        for s, sh, rows, cols, nr, _, i in self._tables(plane):
            k = C.c_uint32()
            self._check(self._L.mars_hip_table_read_ranked(self._ctx, i, C.byref(k), None, None),
                        "table_read_ranked")
            kk = k.value
            out = (N.CCandidate * max(nr * kk, 1))()
            cnt = (C.c_uint8 * max(nr, 1))()
            self._check(self._L.mars_hip_table_read_ranked(self._ctx, i, C.byref(k), out, cnt),
                        "table_read_ranked")
            res[(s, sh)] = (rows, cols, [[out[r * kk + j] for j in range(cnt[r])]
                                         for r in range(nr)])
        return res

    def search_method(self, method: int, row: int, col: int, size: int, shift: int,
                      plane: int = 0):
        """(status, candidate or None, evals) of one retrieval method."""
        out = N.CCandidate()
        n = C.c_uint32()
        ev = C.c_uint64()
        rc = self._L.mars_hip_search_method(self._ctx, plane, method, row, col, size, shift, 1,
                                            C.byref(out), C.byref(n), C.byref(ev))
        return rc, (out if n.value else None), ev.value

    def method_tables(self, method: int, plane: int = 0):
        """{(size, shift): (rows, cols, candidates, evals per range)} of one method."""
        res = {}
        for s, sh, rows, cols, _, _, i in self._tables(plane):
            nr = len(rows)
            cands = (N.CCandidate * max(nr, 1))()
            found = (C.c_uint8 * max(nr, 1))()
            evals = (C.c_uint64 * max(nr, 1))()
            self._check(self._L.mars_hip_table_read_method(self._ctx, i, method, cands, found,
                                                           evals), "table_read_method")
            res[(s, sh)] = (rows, cols, [cands[j] if found[j] else None for j in range(nr)],
                            list(evals[:nr]))
        return res

    def _tables(self, plane: int):
        n = C.c_uint32()
        self._check(self._L.mars_hip_table_count(self._ctx, C.byref(n)), "table_count")
        for i in range(n.value):
            pl = C.c_uint32()
            self._check(self._L.mars_hip_table_plane(self._ctx, i, C.byref(pl)), "table_plane")
            if pl.value != plane:
                continue
            s, sh, nr, P = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint32()
            self._check(self._L.mars_hip_table_info(self._ctx, i, C.byref(s), C.byref(sh),
                                                    C.byref(nr), C.byref(P)), "table_info")
            rows = (C.c_uint32 * max(nr.value, 1))()
            cols = (C.c_uint32 * max(nr.value, 1))()
            self._check(self._L.mars_hip_table_read(self._ctx, i, rows, cols, None, None),
                        "table_read")
            yield s.value, sh.value, list(rows[:nr.value]), list(cols[:nr.value]), nr.value, \
                P.value, i

    def tables(self, plane: int = 0):
        """Tables of one plane: list of (size, shift, rows, cols, candidates[list or None per
        range], n_positions)."""
        n = C.c_uint32()
        self._check(self._L.mars_hip_table_count(self._ctx, C.byref(n)), "table_count")
        res = []
        for i in range(n.value):
            pl = C.c_uint32()
            self._check(self._L.mars_hip_table_plane(self._ctx, i, C.byref(pl)), "table_plane")
            if pl.value != plane:
                continue
            s, sh, nr, P = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint32()
            self._check(self._L.mars_hip_table_info(self._ctx, i, C.byref(s), C.byref(sh),
                                                    C.byref(nr), C.byref(P)), "table_info")
            rows = (C.c_uint32 * max(nr.value, 1))()
            cols = (C.c_uint32 * max(nr.value, 1))()
            cands = (N.CCandidate * max(nr.value, 1))()
            found = (C.c_uint8 * max(nr.value, 1))()
            self._check(self._L.mars_hip_table_read(self._ctx, i, rows, cols, cands, found),
                        "table_read")
            res.append((s.value, sh.value, list(rows[:nr.value]), list(cols[:nr.value]),
                        [cands[j] if found[j] else None for j in range(nr.value)], P.value))
        return res

    def table_scores(self, plane: int = 0):
        """{(size, shift): the search kernel's own winner scores} for one plane."""
        n = C.c_uint32()
        self._check(self._L.mars_hip_table_count(self._ctx, C.byref(n)), "table_count")
        res = {}
        for i in range(n.value):
            pl, s, sh, nr, P = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint32()
            self._check(self._L.mars_hip_table_plane(self._ctx, i, C.byref(pl)), "table_plane")
            if pl.value != plane:
                continue
            self._check(self._L.mars_hip_table_info(self._ctx, i, C.byref(s), C.byref(sh),
                                                    C.byref(nr), C.byref(P)), "table_info")
            out = np.zeros(max(nr.value, 1), np.float64)
            self._check(self._L.mars_hip_table_scores(
                self._ctx, i, out.ctypes.data_as(C.POINTER(C.c_double))), "table_scores")
            res[(s.value, sh.value)] = out[:nr.value]
        return res

    def fit_f64(self, moments: np.ndarray, max_alfa: float, bits_alfa: int, bits_beta: int):
        m = np.ascontiguousarray(moments, dtype=np.int64).reshape(-1, 6)
        n = m.shape[0]
        qa = np.zeros(n, np.uint32)
        qb = np.zeros(n, np.uint32)
        rms = np.zeros(n, np.float64)
        self._check(self._L.mars_hip_fit_f64(
            self._ctx, m.ctypes.data_as(C.POINTER(C.c_int64)), n, max_alfa, bits_alfa, bits_beta,
            qa.ctypes.data_as(C.POINTER(C.c_uint32)), qb.ctypes.data_as(C.POINTER(C.c_uint32)),
            rms.ctypes.data_as(C.POINTER(C.c_double))), "mars_hip_fit_f64")
        return qa, qb, rms

    def copy_probe(self, nbytes: int) -> float:
        ms = C.c_double()
        self._check(self._L.mars_hip_copy_probe(self._ctx, nbytes, C.byref(ms)),
                    "mars_hip_copy_probe")
        return ms.value

    def close(self) -> None:
        if self._ctx:
            self._L.mars_hip_close(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HipSearchProvider:
    """SearchProvider (encode.rs:557-579) backed by the MI355X search table.

    Built once per plane, as encode_image_with_search's factory is
    (encode.rs:1054-1076).  ``shifts`` lists every domain stride the walk
    may request: ``params.shift``, plus ``2 * params.shift`` when adaptive
    density is on (encode.rs:432-437).
    """

    def __init__(self, image: np.ndarray, params: EncodeParams,
                 shifts: Optional[Sequence[int]] = None, device: int = 0,
                 kernel: int = N.KERNEL_MFMA_I8):
        self.params = params
        self.ctx = HipContext(device, kernel)
        self.ctx.prepare(image, params, list(shifts) if shifts else [params.shift])

    def search(self, request: SearchRequest) -> SearchOutcome:
        rc, c, evals = self.ctx.search_raw(request.row, request.col, request.size, request.shift)
        if rc != N.OK:
            msg = self.ctx._L.mars_hip_last_error(self.ctx._ctx).decode(errors="replace")
            raise ValueError(msg)
        return SearchOutcome(_cand(c) if c is not None else None, evals)

    def search_candidates(self, request: SearchRequest, max_: int) -> RankedCandidates:
        """The trait's default (encode.rs:572-578): the winner only."""
        out = self.search(request)
        cands = [out.candidate] if out.candidate is not None else []
        return RankedCandidates(cands[:max(max_, 1)], out.evals)
