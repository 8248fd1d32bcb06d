"""Search providers built from the oracle, for driving the encoder chain in tests."""

import ctypes as C

from mars_amd import _native as N
from mars_amd.codec import SEARCH_FN
from mars_amd.provider import EncodeParams
from tests import _oracle as O


class TableProvider:
    """Oracle winners for every searchable node, served as a mars_search_fn."""

    def __init__(self, img, params: EncodeParams, shifts):
        h, w = img.shape
        self.table = {}
        nodes = O.nodes(w, h, params.min_size, params.max_size)
        for shift in shifts:
            for size, rc in nodes.items():
                rows = [r for r, _ in rc]
                cols = [c for _, c in rc]
                res, _ = O.search_table(img, size, shift, rows, cols, threads=16,
                                        max_alfa=params.max_alfa, bits_alfa=params.bits_alfa,
                                        bits_beta=params.bits_beta)
                P = 0
                if h >= 2 * size and w >= 2 * size:
                    P = ((h - 2 * size) // shift + 1) * ((w - 2 * size) // shift + 1)
                for r, c, cand in zip(rows, cols, res):
                    self.table[(r, c, size, shift)] = (cand, P * 8)
        self.calls = 0

        def cb(user, row, col, size, shift, out, n_out, evals):
            self.calls += 1
            cand, ev = self.table[(row, col, size, shift)]
            evals[0] = ev
            if cand is None:
                n_out[0] = 0
            else:
                C.memmove(out, C.byref(cand), C.sizeof(N.CCandidate))
                n_out[0] = 1
            return 0

        self._cb = SEARCH_FN(cb)
        self.fn = C.cast(self._cb, C.c_void_p)
