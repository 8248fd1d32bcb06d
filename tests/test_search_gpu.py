"""GPU parity of the HIP search against the CPU oracle (bit-exact codes and rms).

Every test calls the product through the C ABI (libmars_hip.so).  The oracle
runs the reference's exhaustive scan on the same plane; parity is equality of
(dom_row, dom_col, isometry, qalfa, qbeta, rms bits, RawMoments) per range.
"""

import struct

import numpy as np
import pytest

from mars_amd import _native as N
from mars_amd import synth
from mars_amd.provider import EncodeParams, HipContext, HipSearchProvider, SearchRequest
from tests import _oracle as O
from tests._checks import winner_consistency

pytestmark = pytest.mark.gpu


def _bits(x: float) -> int:
    return struct.unpack("<q", struct.pack("<d", x))[0]


def _key(c):
    if c is None:
        return None
    m = c.m if hasattr(c, "m") else c
    return (c.dom_row, c.dom_col, c.isometry, c.qalfa, c.qbeta, _bits(c.rms), m.s0, m.s1_x4,
            m.s2_x16, m.t0, m.t1_x4, m.t2)


def _check_tables(img, ctx, label, sample=None, seed=0, **quant):
    h, w = img.shape
    mism = 0
    total = 0
    rng = np.random.default_rng(seed)
    for size, shift, rows, cols, cands, P in ctx.tables():
        idx = np.arange(len(rows))
        if sample is not None and len(idx) > sample:
            idx = np.sort(rng.choice(idx, size=sample, replace=False))
        rr = [rows[i] for i in idx]
        cc = [cols[i] for i in idx]
        ref, _ = O.search_table(img, size, shift, rr, cc, threads=16, **quant)
        for j, i in enumerate(idx):
            total += 1
            if _key(cands[i]) != _key(ref[j]):
                mism += 1
                if mism <= 5:
                    print(f"{label}: size {size} shift {shift} range ({rows[i]},{cols[i]}): "
                          f"gpu {_key(cands[i])} ref {_key(ref[j])}", flush=True)
    assert mism == 0, f"{label}: {mism}/{total} ranges differ"
    return total


def test_device_fit_is_bitwise_binary64():
    ctx = HipContext(0)
    rng = np.random.default_rng(1)
    ms = []
    for size in (2, 4, 8, 16, 32):
        n = size * size
        for _ in range(4000):
            d = rng.integers(0, 1021, size=n).astype(np.int64)
            r = rng.integers(0, 256, size=n).astype(np.int64)
            if rng.random() < 0.3:
                r = np.clip((d * rng.integers(1, 16)) // 64 + rng.integers(-10, 40), 0, 255)
            if rng.random() < 0.05:
                d[:] = d[0]
            ms.append((n, d.sum(), (d * d).sum(), r.sum(), (r * d).sum(), (r * r).sum()))
    ms = np.array(ms, dtype=np.int64)
    qa, qb, rms = ctx.fit_f64(ms, 1.0, 4, 7)
    bad = 0
    for i, m in enumerate(ms):
        eqa, eqb, erms = O.fit_f64(tuple(int(x) for x in m))
        if (qa[i], qb[i], _bits(rms[i])) != (eqa, eqb, _bits(erms)):
            bad += 1
    assert bad == 0, f"{bad} of {len(ms)} device fits differ from the host binary64 fit"


@pytest.mark.parametrize("kernel", [N.KERNEL_MFMA_I8])
def test_parity_random_quadtree_64(kernel):
    img = np.random.default_rng(2).integers(0, 256, size=(64, 64), dtype=np.uint8)
    ctx = HipContext(0, kernel)
    ctx.prepare(img, EncodeParams(min_size=2, max_size=32, shift=4), [4])
    assert _check_tables(img, ctx, "random64") > 0


def test_parity_synthetic_quadtree_128_two_shifts():
    img = synth.plane(128, 128, 21)
    ctx = HipContext(0)
    ctx.prepare(img, EncodeParams(min_size=4, max_size=32, shift=4), [4, 8])
    _check_tables(img, ctx, "synth128")


def test_parity_odd_geometry_border_nodes():
    img = synth.plane(70, 54, 5)
    ctx = HipContext(0)
    ctx.prepare(img, EncodeParams(min_size=4, max_size=16, shift=4), [4])
    sizes = {t[0] for t in ctx.tables()}
    assert 2 in sizes  # forced border splits reach size 2
    _check_tables(img, ctx, "odd70x54")


@pytest.mark.parametrize("shift", [1, 2, 3, 6])
def test_parity_other_strides(shift):
    img = synth.plane(48, 40, 30 + shift)
    ctx = HipContext(0)
    ctx.prepare(img, EncodeParams(min_size=4, max_size=8, shift=shift), [shift])
    _check_tables(img, ctx, f"shift{shift}")


@pytest.mark.parametrize("max_alfa,bits_alfa,bits_beta", [
    (0.5, 4, 7),    # power-of-two scale: the exact integer stage 2 runs
    (2.0, 3, 6),
    (1.0, 8, 12),   # the widest codes stage 2 takes
    (0.75, 4, 7),   # not a power of two: every stage-1 survivor gets the binary64 fit
    (1.0, 9, 7),    # alfa codes beyond stage 2's range: likewise
    (1.0, 4, 13),   # beta codes beyond stage 2's range: likewise
])
def test_parity_quantiser_settings(max_alfa, bits_alfa, bits_beta):
    """The search parameters the fit and its quantisers depend on (encode.rs:247-289)."""
    img = synth.plane(96, 64, 40 + bits_alfa)
    print(f"quantiser max_alfa {max_alfa} bits {bits_alfa}/{bits_beta}", flush=True)
    ctx = HipContext(0)
    ctx.prepare(img, EncodeParams(min_size=4, max_size=16, shift=4, max_alfa=max_alfa,
                                  bits_alfa=bits_alfa, bits_beta=bits_beta), [4])
    _check_tables(img, ctx, f"quant{max_alfa}/{bits_alfa}/{bits_beta}", max_alfa=max_alfa,
                  bits_alfa=bits_alfa, bits_beta=bits_beta)


def test_parity_flat_plane_all_ties():
    img = np.full((64, 64), 93, dtype=np.uint8)
    ctx = HipContext(0)
    ctx.prepare(img, EncodeParams(min_size=4, max_size=16, shift=4), [4])
    _check_tables(img, ctx, "flat")


@pytest.mark.parametrize("kind", ["white", "near_white"])
def test_parity_saturated_planes(kind):
    """Sizes 4 and 8 correlate in binary16 with binary32 accumulation, exact
    while t1 = sum r*D < 2^24.  A white plane puts every t1 at its largest
    value (s = 8: 64 * 255 * 1020 = 16 646 400, 2^24 - 130 816); a plane of
    random values in 240..255 keeps every t1 within 12% of it with the low
    bits varying.  Any rounding in the accumulation shows in t1_x4, the rms
    bits or the winner."""
    if kind == "white":
        img = np.full((48, 48), 255, dtype=np.uint8)
    else:
        img = np.random.default_rng(31).integers(240, 256, size=(72, 64), dtype=np.uint8)
    ctx = HipContext(0)
    ctx.prepare(img, EncodeParams(min_size=4, max_size=8, shift=2), [2])
    assert _check_tables(img, ctx, kind) > 0


def test_parity_fixed_grid_256():
    """256x256 seed-1 plane, fixed 8x8 ranges, stride 8 (the CPU-runnable workload)."""
    img = synth.plane(256, 256, synth.WORKLOAD_SEEDS["fixed256"])
    ctx = HipContext(0)
    ctx.prepare(img, EncodeParams(min_size=8, max_size=8, shift=8), [8])
    assert _check_tables(img, ctx, "fixed256") == 1024


def test_parity_quadtree_512_sampled():
    """512x512 seed-2 plane, quadtree 32->4, stride 4: every size-32/16 range and a
    sample of size-8/4 ranges against the oracle."""
    img = synth.plane(512, 512, synth.WORKLOAD_SEEDS["quadtree512"])
    ctx = HipContext(0)
    ctx.prepare(img, EncodeParams(min_size=4, max_size=32, shift=4), [4])
    tabs = {t[0]: t for t in ctx.tables()}
    assert {s: len(t[2]) for s, t in tabs.items()} == {32: 256, 16: 1024, 8: 4096, 4: 16384}
    _check_tables(img, ctx, "quad512", sample=192, seed=3)


@pytest.mark.timeout(300)
def test_parity_dense_2048_fixed16_sampled():
    """The largest pool: 2048x2048 seed-7 plane, fixed 16x16 grid, stride 1
    (16 384 ranges x 4 068 289 positions x 8); a sample of ranges against the
    oracle, plus every range's winner a real position and the table sized."""
    img = synth.plane(2048, 2048, synth.WORKLOAD_SEEDS["dense2048"])
    print("dense2048 fixed 16x16 stride 1: GPU step", flush=True)
    ctx = HipContext(0)
    ctx.prepare(img, EncodeParams(min_size=16, max_size=16, shift=1), [1])
    (size, shift, rows, cols, cands, P), = ctx.tables()
    assert (size, shift, len(rows), P) == (16, 1, 16384, 4068289)
    assert all(c is not None and c.dom_row <= 2048 - 32 and c.dom_col <= 2048 - 32 for c in cands)
    print("dense2048: oracle sample", flush=True)
    _check_tables(img, ctx, "dense2048", sample=6, seed=5)
    ctx.close()


@pytest.mark.timeout(600)
def test_parity_dense_2048_fixed8_sampled():
    """Config 5's 8x8 grid: the 2048x2048 seed-7 plane, stride 1 (65 536 ranges
    x 4 133 089 positions x 8 isometries).  Every winner is checked for a legal
    position and moments and fit consistent with the pixels; a sample of
    ranges equals the oracle's exhaustive search."""
    img = synth.plane(2048, 2048, synth.WORKLOAD_SEEDS["dense2048"])
    print("dense2048 fixed 8x8 stride 1: GPU step", flush=True)
    ctx = HipContext(0)
    ctx.prepare(img, EncodeParams(min_size=8, max_size=8, shift=1), [1])
    (size, shift, rows, cols, cands, P), = ctx.tables()
    assert (size, shift, len(rows), P) == (8, 1, 65536, 4133089)
    assert all(c is not None for c in cands)
    print("dense2048 8x8: winner consistency", flush=True)
    winner_consistency(img, size, rows, cols, cands)
    print("dense2048 8x8: oracle sample", flush=True)
    _check_tables(img, ctx, "dense2048-8", sample=16, seed=8)
    ctx.close()


def test_provider_lookup_and_errors():
    img = synth.plane(64, 64, 8)
    params = EncodeParams(min_size=4, max_size=16, shift=4)
    prov = HipSearchProvider(img, params)
    out = prov.search(SearchRequest(16, 32, 16, 4))
    ref, ev = O.search_top_k(img, 16, 32, 16, 4)
    assert out.evals == ev == ((64 - 32) // 4 + 1) ** 2 * 8
    c = out.candidate
    assert (c.dom_row, c.dom_col, c.isometry, c.qalfa, c.qbeta, _bits(c.rms)) == \
        (ref[0].dom_row, ref[0].dom_col, ref[0].isometry, ref[0].qalfa, ref[0].qbeta,
         _bits(ref[0].rms))
    with pytest.raises(ValueError):
        prov.search(SearchRequest(1, 1, 4, 4))  # not an aligned node
    rc, _, _ = prov.ctx.search_raw(0, 0, 4, 8)  # stride not prepared
    assert rc == N.E_NOT_IN_TABLE


def test_graph_replays_match_eager():
    img = synth.plane(128, 128, 44)
    ctx = HipContext(0)
    ctx.prepare(img, EncodeParams(min_size=4, max_size=32, shift=4), [4])
    eager = [[_key(c) for c in t[4]] for t in ctx.tables()]
    for _ in range(3):
        ctx.run(iters=3, use_graph=True)
        again = [[_key(c) for c in t[4]] for t in ctx.tables()]
        assert again == eager


def test_fp32_baseline_runs_and_mostly_agrees():
    img = synth.plane(128, 128, 45)
    ctx = HipContext(0, N.KERNEL_FP32)
    ctx.prepare(img, EncodeParams(min_size=4, max_size=32, shift=4), [4])
    diff = tot = 0
    for size, shift, rows, cols, cands, P in ctx.tables():
        ref, _ = O.search_table(img, size, shift, rows, cols, threads=16)
        for a, b in zip(cands, ref):
            tot += 1
            if (a.dom_row, a.dom_col, a.isometry, a.qalfa, a.qbeta) != \
                    (b.dom_row, b.dom_col, b.isometry, b.qalfa, b.qbeta):
                diff += 1
    print(f"fp32 baseline: {diff}/{tot} ranges differ from the binary64 reference", flush=True)
    assert diff <= 0.05 * tot


def test_concurrent_lookups_equal_serial():
    """SearchProvider::search is called from rayon workers in data-dependent
    order (encode.rs:1442-1468, 2033-2051): 16 threads looking up every node of
    a prepared 512x512 table in shuffled orders get the serial answers (ctypes
    releases the GIL around each call, so the lookups overlap)."""
    from concurrent.futures import ThreadPoolExecutor

    img = synth.plane(512, 512, synth.WORKLOAD_SEEDS["quadtree512"])
    ctx = HipContext(0)
    ctx.prepare(img, EncodeParams(min_size=4, max_size=32, shift=4), [4, 8])
    nodes = [(r, c, size, shift) for size, shift, rows, cols, _, _ in ctx.tables()
             for r, c in zip(rows, cols)]
    nodes += [(1, 1, 4, 4), (0, 0, 4, 12)]  # not in the table: error path under threads too

    def look(order):
        out = {}
        for i in order:
            rc, cand, ev = ctx.search_raw(*nodes[i])
            out[i] = (rc, _key(cand), ev)
        return out

    serial = look(range(len(nodes)))
    assert serial[len(nodes) - 1][0] == N.E_NOT_IN_TABLE
    orders = [np.random.default_rng(t).permutation(len(nodes)) for t in range(16)]
    with ThreadPoolExecutor(16) as pool:
        results = list(pool.map(look, orders))
    for res in results:
        assert res == serial
    ctx.close()
