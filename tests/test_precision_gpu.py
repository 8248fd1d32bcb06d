"""Dense-stride precision check at small scale (the 2048x2048 study is
`bench.py --precision`): stride 1, fixed 8x8 and 16x16 grids.

The split-INT8 MFMA table must equal the oracle's bit for bit, and its own
winner score must be the binary64 rms.  The FP32 baseline's own binary32
score of the candidate it picked must be close to that candidate's binary64
rms (the FP32 fit of encode.rs:295-337 on moments accumulated in binary32);
the number of ranges where the FP32 kernel picks other codes is reported.
"""

import numpy as np
import pytest

from mars_amd import _native as N
from mars_amd import synth
from mars_amd.provider import EncodeParams, HipContext
from tests import _oracle as O

pytestmark = pytest.mark.gpu


def _codes(c):
    return None if c is None else (c.dom_row, c.dom_col, c.isometry, c.qalfa, c.qbeta)


def test_dense_stride1_mfma_exact_and_fp32_error_bounded():
    img = synth.plane(192, 192, synth.WORKLOAD_SEEDS["dense2048"])
    params = EncodeParams(min_size=8, max_size=16, shift=1)
    got = {}
    for kern in (N.KERNEL_MFMA_I8, N.KERNEL_FP32):
        ctx = HipContext(0, kern)
        ctx.prepare(img, params, [1])
        got[kern] = ({(t[0], t[1]): t for t in ctx.tables()}, ctx.table_scores())
        ctx.close()
    for key, (size, shift, rows, cols, exact, P) in sorted(got[N.KERNEL_MFMA_I8][0].items()):
        ref, _ = O.search_table(img, size, shift, rows, cols, threads=16)
        for c, r in zip(exact, ref):
            assert (c.dom_row, c.dom_col, c.isometry, c.qalfa, c.qbeta, c.rms, c.t1_x4) == \
                (r.dom_row, r.dom_col, r.isometry, r.qalfa, r.qbeta, r.rms, r.m.t1_x4)
        ms = got[N.KERNEL_MFMA_I8][1][key]
        assert all(s == c.rms for s, c in zip(ms, exact))
        fp = got[N.KERNEL_FP32][0][key][4]
        fs = got[N.KERNEL_FP32][1][key]
        err = [abs(s - c.rms) for s, c in zip(fs, fp)]
        rel = [e / c.rms for e, c in zip(err, fp) if c.rms > 0]
        diff = sum(1 for a, b in zip(exact, fp) if _codes(a) != _codes(b))
        print(f"size {size}: {len(rows)} ranges x {P} positions; FP32 picks other codes for "
              f"{diff}; its binary32 score vs binary64 rms: max abs {max(err):.2e}, max rel "
              f"{max(rel):.2e}", flush=True)
        # binary32 moments (t1 and s2 are inexact above 2^24) and the binary32
        # fit's cancellation in `sum`: a bounded error in pixel-value units
        assert max(err) < 0.25


def test_fp32_scores_are_the_binary32_fit_of_exact_moments():
    """The plain FP32 baseline (encode.rs:295-337 fit_f32 on binary32 moments).
    Up to s = 8 every partial sum of its binary32 correlation is an integer
    below 2^24, so its moments are exact and its own score of each winner must
    equal the oracle's or_fit_f32 of that winner's exact moments bit for bit
    (the table's moments are exact integers).  Above, t1 and S2 round in
    binary32; the deviation there is only reported."""
    img = synth.plane(128, 128, 46)
    ctx = HipContext(0, N.KERNEL_FP32)
    ctx.prepare(img, EncodeParams(min_size=2, max_size=32, shift=2), [2])
    scores = ctx.table_scores()
    exact = 0
    for size, shift, rows, cols, cands, P in ctx.tables():
        sc = scores[(size, shift)]
        dev = 0.0
        for c, s in zip(cands, sc):
            m = (c.s0, c.s1_x4, c.s2_x16, c.t0, c.t1_x4, c.t2)
            _, _, r32 = O.fit_f32(m)
            r32 = float(np.float32(r32))
            if size <= 8:
                assert s == r32, f"size {size}: kernel score {s!r} != fit_f32 {r32!r} of {m}"
                exact += 1
            dev = max(dev, abs(s - r32))
        print(f"size {size}: {len(rows)} winners, max |score - fit_f32(exact)| {dev:.3e}",
              flush=True)
    ctx.close()
    assert exact > 1000
