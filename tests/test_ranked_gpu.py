"""GPU parity of ranked (top-k) tables: ExhaustiveSearch::search_candidates.

A table prepared with mars_hip_set_ranked(k) holds each range's best k by
(rms, enumeration order) -- the reference's search_top_k
(crates/mars-codec/src/encode.rs:676-685, 695-788) -- with every record's
moments and fit bit-identical to the oracle's; rank 0 is the k = 1 winner.
The RD walk with rd_candidates > 1 (P5c, encode.rs:1946-1975) then gives
the same `.mars` bytes as the oracle chain.
"""

import pytest

from mars_amd import _native as N
from mars_amd import synth
from mars_amd.codec import EncodeOptions, encode_plane, hip_encode_plane
from mars_amd.provider import EncodeParams, HipContext
from tests import _oracle as O
from tests._providers import TableProvider

pytestmark = pytest.mark.gpu


def _key(c):
    m = c.m if hasattr(c, "m") else c
    return (c.dom_row, c.dom_col, c.isometry, c.qalfa, c.qbeta, c.rms.hex(), m.s0, m.s1_x4,
            m.s2_x16, m.t0, m.t1_x4, m.t2)


@pytest.mark.parametrize("w,h,seed,mn,mx,shifts,k", [
    (96, 64, 11, 4, 16, [4], 4),
    (160, 96, 13, 4, 32, [4], 32),
    (75, 53, 15, 2, 16, [4, 8], 3),   # odd sides, border nodes down to size 2
    (20, 12, 14, 4, 8, [2], 8),       # tiny pools: fewer positions x isometries than k
    (64, 64, 12, 4, 8, [1], 5),       # dense stride
])
def test_ranked_tables_equal_the_oracle(w, h, seed, mn, mx, shifts, k):
    img = synth.plane(w, h, seed)
    print(f"ranked k={k}: {w}x{h} seed {seed} sizes {mn}..{mx} shifts {shifts}", flush=True)
    params = EncodeParams(min_size=mn, max_size=mx, shift=shifts[0])
    ctx = HipContext(0)
    ctx.set_ranked(k)
    ctx.prepare(img, params, shifts)
    one = HipContext(0)
    one.prepare(img, params, shifts)
    winners = {(s, sh): [(_key(c) if c is not None else None) for c in cands]
               for s, sh, rows, cols, cands, _ in one.tables()}
    n = 0
    for (size, shift), (rows, cols, lists) in ctx.ranked_tables().items():
        ref, _ = O.search_table_k(img, size, shift, rows, cols, k, threads=16)
        for i, (got, want) in enumerate(zip(lists, ref)):
            assert [_key(c) for c in got] == [_key(c) for c in want], \
                (size, shift, rows[i], cols[i], len(got), len(want))
            first = _key(got[0]) if got else None
            assert first == winners[(size, shift)][i]
            n += 1
    assert n > 0
    ctx.close()
    one.close()


def test_ranked_lookup_and_errors():
    img = synth.plane(96, 64, 11)
    ctx = HipContext(0)
    with pytest.raises(RuntimeError):
        ctx.set_ranked(33)
    ctx.set_ranked(4)
    ctx.prepare(img, EncodeParams(min_size=4, max_size=16, shift=4), [4])
    rc, got, ev = ctx.search_ranked(0, 0, 8, 4, 3)
    ref, ref_ev = O.search_top_k(img, 0, 0, 8, 4, k=3)
    assert rc == N.OK and ev == ref_ev and [_key(c) for c in got] == [_key(c) for c in ref]
    rc, _, _ = ctx.search_ranked(0, 0, 8, 4, 5)
    assert rc == N.E_UNSUPPORTED  # more than the table ranks
    ctx.set_kernel(N.KERNEL_FP32)
    with pytest.raises(RuntimeError):
        ctx.prepare(img, EncodeParams(min_size=4, max_size=16, shift=4), [4])
    ctx.close()


def test_ranked_graph_replays_keep_the_lists():
    imgs = [synth.plane(96, 64, 11), synth.plane(64, 64, 12)]
    ctx = HipContext(0)
    ctx.set_ranked(6)
    ctx.prepare_batch(imgs, EncodeParams(min_size=4, max_size=16, shift=4), [4])
    before = [{k: [[_key(c) for c in l] for l in v[2]] for k, v in ctx.ranked_tables(i).items()}
              for i in range(2)]
    ctx.run(iters=3, use_graph=True)
    after = [{k: [[_key(c) for c in l] for l in v[2]] for k, v in ctx.ranked_tables(i).items()}
             for i in range(2)]
    assert after == before
    ctx.close()


@pytest.mark.parametrize("rd", [3, 8])
def test_rd_candidates_encode_equals_the_oracle_chain(rd):
    img = synth.plane(160, 96, 77)
    p = EncodeParams(min_size=4, max_size=16, shift=4, lambda_=200.0)
    o = EncodeOptions(allowed_modes=(True, False, True, True), rd_candidates=rd)
    prov = TableProvider(img, p, [4], k=rd)
    ref, ref_st = encode_plane(img, p, o, prov.fn, None)
    ctx = HipContext(0)
    got, st = hip_encode_plane(ctx, img, p, o)
    ctx.close()
    assert (st.search_evals, st.warmup_evals) == (ref_st.search_evals, ref_st.warmup_evals)
    assert got == ref
