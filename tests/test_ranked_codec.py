"""CPU checks of the RD walk with several ranked candidates (P5c).

The reference's own properties (crates/mars-codec/src/encode.rs:3089-3125):
an explicit rd_candidates = 1 is the default encode, and widening the list is
a no-op when only mode 0 is allowed.  Also: the oracle's ranked provider
callback (search_top_k with max) and a precomputed top-k table give the same
bytes, and the search_table_k rows equal search_top_k range by range.
"""

import pytest

from mars_amd import synth
from mars_amd.codec import EncodeOptions, encode_plane
from mars_amd.provider import EncodeParams
from tests import _oracle as O
from tests._providers import TableProvider


def _img():
    return synth.plane(48, 40, 31)


def _enc(img, p, o):
    prov = O.Provider(img, p.max_alfa, p.bits_alfa, p.bits_beta)
    try:
        return encode_plane(img, p, o, prov.fn, prov.user)
    finally:
        prov.close()


def test_explicit_one_is_the_default():
    img = _img()
    p = EncodeParams(min_size=4, max_size=16, shift=4, lambda_=200.0)
    base = EncodeOptions(allowed_modes=(True, False, True, True))
    one = EncodeOptions(allowed_modes=(True, False, True, True), rd_candidates=1)
    assert _enc(img, p, base)[0] == _enc(img, p, one)[0]


def test_mode0_only_ignores_the_candidate_list():
    img = _img()
    p = EncodeParams(min_size=4, max_size=16, shift=4, lambda_=200.0)
    m0 = EncodeOptions(allowed_modes=(True, False, False, False))
    wide = EncodeOptions(allowed_modes=(True, False, False, False), rd_candidates=4)
    assert _enc(img, p, m0)[0] == _enc(img, p, wide)[0]


@pytest.mark.parametrize("rd", [2, 5])
def test_ranked_callback_equals_ranked_table(rd):
    img = _img()
    p = EncodeParams(min_size=4, max_size=16, shift=4, lambda_=50.0)
    o = EncodeOptions(allowed_modes=(True, True, True, True), rd_candidates=rd)
    a, sa = _enc(img, p, o)
    tab = TableProvider(img, p, [4], k=rd)
    b, sb = encode_plane(img, p, o, tab.fn, None)
    assert a == b and sa.search_evals == sb.search_evals
    assert _enc(img, p, o)[0] == a  # deterministic


def test_search_table_k_rows_equal_search_top_k():
    img = _img()
    rows, cols = [0, 8, 16, 24], [0, 16, 8, 32]
    lists, _ = O.search_table_k(img, 8, 4, rows, cols, 6, threads=3)
    for r, c, got in zip(rows, cols, lists):
        want, _ = O.search_top_k(img, r, c, 8, 4, k=6)
        assert [(x.dom_row, x.dom_col, x.isometry, x.rms) for x in got] == \
               [(x.dom_row, x.dom_col, x.isometry, x.rms) for x in want]
        assert all(a.rms <= b.rms for a, b in zip(got, got[1:]))
