"""GPU parity of the top-32 oracle cache (crates/mars-bench/src/oracle.rs:539-575).

The cache built from one ranked (k = 32) step on the GPU is byte-identical
to the cache assembled from the oracle's search_top_k, and it scores the
GPU's own k = 1 winners as perfect top-1 recall with zero regret.
"""

import pytest

from mars_amd import oracle_cache as OC
from mars_amd import synth
from mars_amd.provider import EncodeParams, HipContext
from tests._providers import oracle_cache

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("w,h,seed,mn,mx,shift", [
    (96, 64, 21, 4, 16, 4),
    (75, 53, 22, 4, 32, 8),   # odd sides: grid blocks stop short of the border
])
def test_gpu_cache_bytes_equal_the_oracle_cache(w, h, seed, mn, mx, shift):
    img = synth.plane(w, h, seed)
    config = OC.OracleConfig(mn, mx, shift)
    print(f"oracle cache: {w}x{h} seed {seed} sizes {mn}..{mx} shift {shift}", flush=True)
    ctx = HipContext(0)
    got = OC.build(ctx, img, "synthetic", config)
    ctx.close()
    want = oracle_cache(img, config)
    assert OC.serialize(got) == OC.serialize(want)


def test_gpu_winners_score_as_top1():
    img = synth.plane(96, 64, 23)
    config = OC.OracleConfig(4, 16, 4)
    ctx = HipContext(0)
    cache = OC.build(ctx, img, "synthetic", config)
    one = HipContext(0)
    one.prepare(img, EncodeParams(min_size=4, max_size=16, shift=4), [4])
    picks = []
    for size, sh, rows, cols, cands, _ in one.tables():
        for r, c, cand in zip(rows, cols, cands):
            if cand is not None and r + size <= 64 and c + size <= 96:
                picks.append(OC.MethodPick(r, c, size, cand.dom_row, cand.dom_col, cand.isometry,
                                           cand.qalfa, cand.qbeta, cand.rms))
    one.close()
    ctx.close()
    rep = OC.score(cache, picks)
    assert rep.matched == len(picks) == sum(sc.num_blocks_x * sc.num_blocks_y
                                            for sc in cache.sizes)
    assert rep.top1 == rep.matched
