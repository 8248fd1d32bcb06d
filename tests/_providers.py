"""Search providers built from the oracle, for driving the encoder chain in tests."""

import ctypes as C

from mars_amd import _native as N
from mars_amd import oracle_cache as OC
from mars_amd.codec import SEARCH_FN
from mars_amd.provider import EncodeParams
from tests import _oracle as O


class TableProvider:
    """Oracle results for every searchable node, served as a mars_search_fn.

    With k > 1 the table holds each node's best k (search_top_k) and the
    callback returns the first `max` of them.
    """

    def __init__(self, img, params: EncodeParams, shifts, k: int = 1):
        h, w = img.shape
        self.table = {}
        nodes = O.nodes(w, h, params.min_size, params.max_size)
        for shift in shifts:
            for size, rc in nodes.items():
                rows = [r for r, _ in rc]
                cols = [c for _, c in rc]
                res, _ = O.search_table_k(img, size, shift, rows, cols, k, threads=16,
                                          max_alfa=params.max_alfa,
                                          bits_alfa=params.bits_alfa,
                                          bits_beta=params.bits_beta)
                P = 0
                if h >= 2 * size and w >= 2 * size:
                    P = ((h - 2 * size) // shift + 1) * ((w - 2 * size) // shift + 1)
                for r, c, cands in zip(rows, cols, res):
                    self.table[(r, c, size, shift)] = (cands, P * 8)
        self.calls = 0

        def cb(user, row, col, size, shift, mx, out, n_out, evals):
            self.calls += 1
            cands, ev = self.table[(row, col, size, shift)]
            evals[0] = ev
            take = cands[:max(1, mx)]
            for i, cand in enumerate(take):
                C.memmove(C.byref(out[i]), C.byref(cand), C.sizeof(N.CCandidate))
            n_out[0] = len(take)
            return 0

        self._cb = SEARCH_FN(cb)
        self.fn = C.cast(self._cb, C.c_void_p)


def oracle_cache(img, config):
    """The top-32 oracle cache assembled from the oracle's search_top_k (k = 32)."""
    h, w = img.shape
    memo = {}

    def lookup(row, col, size):
        if size not in memo:
            grid = [(r, c) for r in range(0, h - size + 1, size)
                    for c in range(0, w - size + 1, size)]
            lists, _ = O.search_table_k(img, size, config.shift, [r for r, _ in grid],
                                        [c for _, c in grid], OC.SLOTS, threads=16,
                                        max_alfa=config.max_alfa, bits_alfa=config.bits_alfa,
                                        bits_beta=config.bits_beta)
            memo[size] = dict(zip(grid, lists))
        return memo[size][(row, col)]

    return OC.assemble(img, "synthetic", config, lookup)


def _bits(x: float) -> int:
    return C.c_int64.from_buffer_copy(C.c_double(x)).value


def candidate_key(c):
    """Every field of a winner: codes, rms bits and the six RawMoments."""
    if c is None:
        return None
    m = c.m if hasattr(c, "m") else c  # oracle candidates keep RawMoments in .m
    return (c.dom_row, c.dom_col, c.isometry, c.qalfa, c.qbeta, _bits(c.rms), m.s0, m.s1_x4,
            m.s2_x16, m.t0, m.t1_x4, m.t2)


def assert_tables_equal(tables, provider: TableProvider, label: str) -> int:
    """Every (size, shift) table of a context (HipContext.tables() rows) equals the
    oracle table a TableProvider computed, range by range; returns the ranges
    compared.  Every node the provider holds must be in the context's tables."""
    seen = set()
    n = 0
    for size, shift, rows, cols, cands, _ in tables:
        for r, c, got in zip(rows, cols, cands):
            ref, _ = provider.table[(r, c, size, shift)]
            want = ref[0] if ref else None
            assert candidate_key(got) == candidate_key(want), \
                f"{label}: size {size} shift {shift} range ({r},{c}): " \
                f"{candidate_key(got)} != {candidate_key(want)}"
            seen.add((r, c, size, shift))
            n += 1
    assert seen == set(provider.table), f"{label}: the context misses oracle nodes"
    return n
