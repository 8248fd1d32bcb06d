"""GPU parity of the batched step: many planes, one search launch per size.

A batch must give every plane exactly the table that plane gets when it is
prepared alone (and that the oracle gives), whatever the chunking shape the
autotuner picks; encoding a plane from the batch must give the same `.mars`
bytes as the single-plane encode.
"""

import struct

import numpy as np
import pytest

from mars_amd import _native as N
from mars_amd import synth
from mars_amd.codec import EncodeOptions, hip_encode_plane, hip_encode_prepared
from mars_amd.provider import EncodeParams, HipContext
from tests import _oracle as O

pytestmark = pytest.mark.gpu

# Uneven sizes, one too small for 16x16 domains and one with odd sides.
SHAPES = [(96, 64, 11), (64, 64, 12), (160, 96, 13), (20, 12, 14), (75, 53, 15)]


def _bits(x: float) -> int:
    return struct.unpack("<q", struct.pack("<d", x))[0]


def _key(c):
    if c is None:
        return None
    m = c.m if hasattr(c, "m") else c  # oracle candidates keep RawMoments in .m
    return (c.dom_row, c.dom_col, c.isometry, c.qalfa, c.qbeta, _bits(c.rms), m.s0, m.s1_x4,
            m.s2_x16, m.t0, m.t1_x4, m.t2)


def _planes():
    return [synth.plane(w, h, seed) for w, h, seed in SHAPES]


def _table_keys(ctx, plane):
    return {(s, sh): (rows, cols, [_key(c) for c in cands], P)
            for s, sh, rows, cols, cands, P in ctx.tables(plane)}


def test_batch_equals_single_plane_and_oracle():
    imgs = _planes()
    params = EncodeParams(min_size=4, max_size=16, shift=4)
    batch = HipContext(0)
    batch.prepare_batch(imgs, params, [4, 8])
    for i, img in enumerate(imgs):
        print(f"plane {i} {img.shape}", flush=True)
        one = HipContext(0)
        one.prepare(img, params, [4, 8])
        got, ref = _table_keys(batch, i), _table_keys(one, 0)
        assert got == ref, f"plane {i}: batch tables differ from the single-plane tables"
        one.close()
        for (size, shift), (rows, cols, keys, P) in got.items():
            oref, _ = O.search_table(img, size, shift, rows, cols, threads=16)
            assert keys == [_key(c) for c in oref], f"plane {i} size {size} shift {shift}"
    batch.close()


@pytest.mark.parametrize("kernel", [N.KERNEL_MFMA_I8, N.KERNEL_FP32])
def test_autotuned_shapes_keep_results(kernel):
    imgs = _planes()[:3]
    params = EncodeParams(min_size=4, max_size=16, shift=4)
    ctx = HipContext(0, kernel)
    ctx.prepare_batch(imgs, params, [4])
    before = [_table_keys(ctx, i) for i in range(len(imgs))]
    t = ctx.autotune(reps=1)
    print("tuning", list(t.target_blocks), list(t.min_tiles), flush=True)
    after = [_table_keys(ctx, i) for i in range(len(imgs))]
    assert after == before
    ctx.run(iters=2, use_graph=True)
    assert [_table_keys(ctx, i) for i in range(len(imgs))] == before
    ctx.close()


def test_encode_from_batch_matches_single_plane_encode():
    imgs = _planes()[:3]
    params = EncodeParams(min_size=4, max_size=16, shift=4, lambda_=200.0)
    options = EncodeOptions(allowed_modes=(True, False, True, False))
    batch = HipContext(0)
    batch.prepare_batch(imgs, params, [4])
    for i, img in enumerate(imgs):
        got, st = hip_encode_prepared(batch, i, params, options)
        one = HipContext(0)
        ref, ref_st = hip_encode_plane(one, img, params, options)
        one.close()
        assert (st.search_evals, st.warmup_evals) == (ref_st.search_evals, ref_st.warmup_evals)
        assert got == ref, f"plane {i}"
    batch.close()


def test_batch_lookup_errors():
    imgs = _planes()[:2]
    ctx = HipContext(0)
    ctx.prepare_batch(imgs, EncodeParams(min_size=4, max_size=16, shift=4), [4])
    rc, c, _ = ctx.search_raw(0, 0, 4, 4, plane=1)
    assert rc == N.OK and c is not None
    rc, _, _ = ctx.search_raw(0, 0, 4, 4, plane=2)
    assert rc == N.E_NOT_IN_TABLE
    with pytest.raises(RuntimeError):
        hip_encode_prepared(ctx, 0, EncodeParams(min_size=8, max_size=16, shift=4),
                            EncodeOptions())
    ctx.close()
