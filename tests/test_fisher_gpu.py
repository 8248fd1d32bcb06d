"""GPU parity of the Fisher classified search and the colour encode.

* Every Fisher table (winner, its moments, fit and evals of every range) is
  bit-identical to the oracle's search_block_fitted over the range's bucket,
  and preparing with Fisher leaves the exhaustive tables unchanged.
* `.mars` bytes: a plane encoded with the GPU Fisher provider equals the
  same chain with the oracle's Fisher provider (RD warm-up exhaustive on
  both sides, as in the reference); the colour encode (BT.601, 4:2:0, MARC)
  equals its oracle composition, at the 768x512 RGB workload too.
"""

import hashlib

import numpy as np
import pytest

from mars_amd import _native as N
from mars_amd import codec, synth
from mars_amd.codec import EncodeOptions, encode_plane, hip_encode_plane_method, hip_encode_rgb
from mars_amd.provider import EncodeParams, HipContext
from tests import _oracle as O
from tests._providers import TableProvider, assert_tables_equal

pytestmark = pytest.mark.gpu

FISHER = N.METHOD_FISHER


def _key(c):
    if c is None:
        return None
    m = c.m if hasattr(c, "m") else c
    return (c.dom_row, c.dom_col, c.isometry, c.qalfa, c.qbeta, c.rms.hex(), m.s0, m.s1_x4,
            m.s2_x16, m.t0, m.t1_x4, m.t2)


def _check_tables(ctx, img, plane=0):
    n = 0
    for (size, shift), (rows, cols, cands, evals) in ctx.method_tables(FISHER, plane).items():
        ref, ref_evals = O.fisher_table(img, size, shift, rows, cols, threads=16)
        got = [_key(c) for c in cands]
        want = [_key(c) for c in ref]
        bad = [i for i in range(len(rows)) if got[i] != want[i] or evals[i] != ref_evals[i]]
        assert not bad, (f"size {size} shift {shift}: {len(bad)} of {len(rows)} ranges differ, "
                         f"first ({rows[bad[0]]},{cols[bad[0]]}): {got[bad[0]]} vs {want[bad[0]]}, "
                         f"evals {evals[bad[0]]} vs {ref_evals[bad[0]]}")
        n += len(rows)
    return n


@pytest.mark.parametrize("w,h,seed,mn,mx,shifts", [
    (96, 64, 11, 4, 16, [4]),
    (75, 53, 15, 2, 16, [4, 8]),   # odd sides: border nodes down to size 2
    (20, 12, 14, 4, 8, [2]),       # positions only for the smallest sizes
    (160, 96, 13, 4, 32, [4]),
    (64, 64, 12, 4, 8, [1]),       # dense stride
])
def test_fisher_tables_equal_the_oracle(w, h, seed, mn, mx, shifts):
    img = synth.plane(w, h, seed)
    print(f"fisher tables {w}x{h} seed {seed} sizes {mn}..{mx} shifts {shifts}", flush=True)
    ctx = HipContext(0)
    ctx.prepare(img, EncodeParams(min_size=mn, max_size=mx, shift=shifts[0]), shifts,
                method=FISHER)
    assert _check_tables(ctx, img) > 0
    # the exhaustive tables of the same context still equal the oracle
    for size, shift, rows, cols, cands, _ in ctx.tables():
        ref, _ = O.search_table(img, size, shift, rows, cols, threads=16)
        assert [_key(c) for c in cands] == [_key(c) for c in ref], (size, shift)
    ctx.close()


def test_fisher_lookup_and_errors():
    img = synth.plane(96, 64, 11)
    ctx = HipContext(0)
    ctx.prepare(img, EncodeParams(min_size=4, max_size=16, shift=4), [4], method=FISHER)
    rc, c, ev = ctx.search_raw(0, 0, 8, 4)  # the prepared method: Fisher
    rc2, c2, ev2 = ctx.search_method(FISHER, 0, 0, 8, 4)
    assert rc == rc2 == N.OK and _key(c) == _key(c2) and ev == ev2
    ref, ref_ev = O.fisher_table(img, 8, 4, [0], [0])
    assert _key(c) == _key(ref[0]) and ev == ref_ev[0]
    rc, c, ev = ctx.search_method(N.METHOD_EXHAUSTIVE, 0, 0, 8, 4)
    assert rc == N.OK and ev == ((64 - 16) // 4 + 1) * ((96 - 16) // 4 + 1) * 8
    ctx.prepare(img, EncodeParams(min_size=4, max_size=16, shift=4), [4])
    rc, _, _ = ctx.search_method(FISHER, 0, 0, 8, 4)
    assert rc == N.E_ARG  # Fisher was not prepared
    ctx.close()


def test_fisher_batch_equals_single_planes():
    imgs = [synth.plane(96, 64, 11), synth.plane(75, 53, 15), synth.plane(64, 64, 12)]
    params = EncodeParams(min_size=4, max_size=16, shift=4)
    batch = HipContext(0)
    batch.prepare_batch(imgs, params, [4], method=FISHER)
    for i, img in enumerate(imgs):
        print(f"batch plane {i}", flush=True)
        _check_tables(batch, img, plane=i)
    batch.run(iters=3, use_graph=True)  # graph replays keep the tables
    for i, img in enumerate(imgs):
        _check_tables(batch, img, plane=i)
    batch.close()


def _oracle_encode(img, p, o, shifts, keep=None):
    fp = O.FisherProvider(img, p.max_alfa, p.bits_alfa, p.bits_beta)
    warm = TableProvider(img, p, shifts) if p.lambda_ is not None else None
    data, st = encode_plane(img, p, o, fp.fn, fp.user, warm.fn if warm else None, None)
    fp.close()
    if keep is not None:
        keep.append((img, p, warm))
    return data, st


@pytest.mark.parametrize("lam,modes,density,tr", [
    (None, (True, True, True, True), False, 8.0),
    (200.0, (True, False, True, False), False, 8.0),
    (50.0, (True, True, True, True), True, 8.0),
])
def test_fisher_encode_equals_the_oracle_chain(lam, modes, density, tr):
    img = synth.plane(160, 96, 77)
    p = EncodeParams(min_size=4, max_size=16, shift=4, lambda_=lam, t_rms=tr)
    o = EncodeOptions(allowed_modes=modes, adaptive_density=density,
                      lambda_adaptive_qstep=density)
    shifts = [4, 8] if density and lam is not None else [4]
    ref, ref_st = _oracle_encode(img, p, o, shifts)
    ctx = HipContext(0)
    got, st = hip_encode_plane_method(ctx, img, p, o, FISHER)
    ctx.close()
    assert (st.search_evals, st.warmup_evals) == (ref_st.search_evals, ref_st.warmup_evals)
    assert got == ref


def _oracle_rgb(rgb, yp, cp, o, keep=None):
    y, cb, cr = O.ycbcr(rgb)
    streams, stats = [], []
    for plane, p in ((y, yp), (O.downsample_box(cb), cp), (O.downsample_box(cr), cp)):
        data, st = _oracle_encode(plane, p, o, [p.shift], keep)
        streams.append(data)
        stats.append(st)
    return codec.write_container(codec.COLOR_RGB_420, streams), stats


def _check_every_winner(planes):
    """Every winner of both methods on every plane of a colour encode: the
    exhaustive tables (full, not the walk-only ones the encode reads) against
    the oracle tables the warm-up provider computed, the Fisher tables against
    the oracle's search_block_fitted.  The bytes alone would miss a wrong
    winner that changes no decision."""
    for name, (img, p, warm) in zip(("y", "cb", "cr"), planes):
        ctx = HipContext(0)
        ctx.prepare(img, p, [p.shift], method=FISHER)
        n = assert_tables_equal(ctx.tables(), warm, name)
        f = _check_tables(ctx, img)
        ctx.close()
        print(f"  {name}: {n} exhaustive and {f} Fisher winners equal the oracle's", flush=True)


@pytest.mark.parametrize("w,h", [
    (192, 128),
    # the oracle's exhaustive warm-up table of the 768x512 Y plane takes about
    # a minute of all host cores
    pytest.param(768, 512, marks=pytest.mark.timeout(900)),
])
def test_rgb420_fisher_encode_equals_the_oracle_chain(w, h):
    """The colour workload: RGB -> YCbCr 4:2:0, Y 32->4, Cb/Cr 16->4, lambda 200, Fisher."""
    rgb = synth.rgb(w, h, synth.WORKLOAD_SEEDS["rgb768x512"])
    yp = EncodeParams(min_size=4, max_size=32, shift=4, lambda_=200.0)
    cp = EncodeParams(min_size=4, max_size=16, shift=4, lambda_=200.0)
    o = EncodeOptions(allowed_modes=(True, False, True, False))
    print(f"rgb {w}x{h}: oracle chain", flush=True)
    planes = []
    ref, ref_stats = _oracle_rgb(rgb, yp, cp, o, planes)
    print("rgb: GPU", flush=True)
    ctx = HipContext(0)
    got, stats, nbytes = hip_encode_rgb(ctx, rgb, yp, cp, o, FISHER)
    ctx.close()
    print(f"rgb {w}x{h}: {len(got)} bytes sha256 {hashlib.sha256(got).hexdigest()} "
          f"planes {nbytes} evals {[s.search_evals for s in stats.values()]}", flush=True)
    for k, rs in zip(("y", "cb", "cr"), ref_stats):
        assert (stats[k].search_evals, stats[k].warmup_evals) == (rs.search_evals,
                                                                  rs.warmup_evals), k
    assert got == ref
    _check_every_winner(planes)


def test_rgb444_exhaustive_equals_the_oracle_chain():
    """Method exhaustive through the colour path: encode_color_image's own search."""
    rgb = synth.rgb(96, 64, 5)
    p = EncodeParams(min_size=4, max_size=16, shift=4, lambda_=None)
    o = EncodeOptions()
    ctx = HipContext(0)
    got, _, _ = hip_encode_rgb(ctx, rgb, p, p, o, N.METHOD_EXHAUSTIVE,
                               subsampling=codec.COLOR_RGB_444)
    ctx.close()
    streams = []
    for plane in O.ycbcr(rgb):
        prov = TableProvider(plane, p, [4])
        data, _ = encode_plane(plane, p, o, prov.fn, None)
        streams.append(data)
    assert got == codec.write_container(codec.COLOR_RGB_444, streams)
    assert np.frombuffer(got[:7], np.uint8).tolist()[:7] == list(b"MARC") + [0, 1, 3]
