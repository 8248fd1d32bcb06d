"""GPU parity of the uneven 64-image batch workload (synth.batch_specs).

The whole batch -- 64 planes drawn from {256x256, 512x512, 768x512, 1024x768},
quadtree 32->4, stride 4 -- is prepared as one batch, as the bench runs it.
Paths that only a batch of this size reaches (per-group work-list offsets
over 64 planes, the chunk-count clamp of the partial buffers, arena sizing)
are covered here:

* every plane's batched tables equal the tables of that plane prepared
  alone, before and after autotuning the chunking shapes on the batch;
* 16 sampled ranges per (plane, size) equal the oracle's exhaustive search
  (encode.rs:695-788);
* one 256x256 and one 512x512 plane, encoded from the batch at their drawn
  lambda (modes {0, 2}), give the same `.mars` bytes as the oracle chain.
"""

import numpy as np
import pytest

from mars_amd import synth
from mars_amd.codec import EncodeOptions, encode_plane, hip_encode_prepared
from mars_amd.provider import EncodeParams, HipContext
from tests import _oracle as O
from tests._providers import TableProvider, candidate_key

pytestmark = pytest.mark.gpu

PARAMS = EncodeParams(min_size=4, max_size=32, shift=4)
SAMPLE = 16


def _keys(ctx, plane):
    return {(s, sh): (rows, cols, [candidate_key(c) for c in cands], P)
            for s, sh, rows, cols, cands, P in ctx.tables(plane)}


@pytest.fixture(scope="module")
def batch():
    specs = synth.batch_specs()
    imgs = [synth.plane(w, h, seed) for w, h, seed, _ in specs]
    print(f"batch64: preparing {len(imgs)} planes in one batch", flush=True)
    ctx = HipContext(0)
    ctx.prepare_batch(imgs, PARAMS, [PARAMS.shift])
    yield specs, imgs, ctx
    ctx.close()


@pytest.mark.timeout(600)
def test_batch64_equals_single_planes_and_oracle(batch):
    specs, imgs, ctx = batch
    first = [_keys(ctx, i) for i in range(len(imgs))]
    rng = np.random.default_rng(64)
    checked = 0
    for i, (img, (w, h, seed, _)) in enumerate(zip(imgs, specs)):
        one = HipContext(0)
        one.prepare(img, PARAMS, [PARAMS.shift])
        alone = _keys(one, 0)
        one.close()
        assert first[i] == alone, f"plane {i} ({w}x{h} seed {seed}): batch != single plane"
        for (size, shift), (rows, cols, keys, P) in sorted(first[i].items()):
            idx = np.sort(rng.choice(len(rows), size=min(SAMPLE, len(rows)), replace=False))
            ref, _ = O.search_table(img, size, shift, [rows[j] for j in idx],
                                    [cols[j] for j in idx], threads=16)
            for j, r in zip(idx, ref):
                assert keys[j] == candidate_key(r), \
                    f"plane {i} size {size} range ({rows[j]},{cols[j]}): {keys[j]} vs oracle"
                checked += 1
        print(f"plane {i} {w}x{h}: tables = single plane = oracle sample", flush=True)
    print(f"batch64: {checked} sampled ranges equal the oracle", flush=True)
    # the shapes the autotuner picks on this batch (the bench's) keep every table
    t = ctx.autotune(reps=1)
    print("batch64 tuning", list(t.target_blocks), list(t.min_tiles), flush=True)
    ctx.run(iters=2, use_graph=True)
    assert [_keys(ctx, i) for i in range(len(imgs))] == first


@pytest.mark.timeout(600)
@pytest.mark.parametrize("want", [(256, 256), (512, 512)])
def test_batch64_encode_equals_the_oracle_chain(batch, want):
    specs, imgs, ctx = batch
    i = next(k for k, (w, h, _, _) in enumerate(specs) if (w, h) == want)
    w, h, seed, lam = specs[i]
    p = EncodeParams(min_size=4, max_size=32, shift=4, lambda_=lam)
    o = EncodeOptions(allowed_modes=(True, False, True, False))
    print(f"batch64 plane {i} {w}x{h} seed {seed} lambda {lam}: oracle chain", flush=True)
    prov = TableProvider(imgs[i], p, [4])
    ref, ref_st = encode_plane(imgs[i], p, o, prov.fn, None)
    got, st = hip_encode_prepared(ctx, i, p, o)
    assert (st.search_evals, st.warmup_evals) == (ref_st.search_evals, ref_st.warmup_evals)
    assert got == ref, f"plane {i}: {len(got)} bytes vs the oracle chain's {len(ref)}"
