"""Walk-only exhaustive tables (mars_hip_set_walk_tables) on the GPU.

A Fisher batch whose exhaustive tables only serve the RD warm-up computes,
size by size on the device, the winners of exactly the nodes that warm-up's
threshold walk visits (encode.rs:1330-1435: a searched node splits when its
best rms exceeds t_rms -- +inf without a candidate -- and its size exceeds
min_size; nodes larger than max_size or across the border split without a
search).  The visited set is recomputed here by walking the
full table; every visited node's winner must equal the full table's, every
other exhaustive lookup must fail, and the Fisher tables must not change.
"""

import math

import pytest

from mars_amd import _native as N
from mars_amd import synth
from mars_amd.provider import EncodeParams, HipContext

pytestmark = pytest.mark.gpu

EXH, FISHER = N.METHOD_EXHAUSTIVE, N.METHOD_FISHER


def _key(c):
    return None if c is None else (c.dom_row, c.dom_col, c.isometry, c.qalfa, c.qbeta,
                                   c.rms.hex(), c.s1_x4, c.s2_x16, c.t0, c.t1_x4, c.t2)


def _visited(w, h, p, shift, t_rms, full):
    out = set()

    def walk(row, col, size):
        if row >= h or col >= w:
            return
        if size > p.max_size or row + size > h or col + size > w:
            half = size // 2
            for r, c in ((row, col), (row + half, col), (row, col + half), (row + half, col + half)):
                walk(r, c, half)
            return
        if size == 1:
            return
        out.add((size, row, col))
        c = full[(size, shift)][(row, col)]
        rms = c.rms if c is not None else math.inf
        if rms > t_rms and size > p.min_size:
            half = size // 2
            for r, c in ((row, col), (row + half, col), (row, col + half), (row + half, col + half)):
                walk(r, c, half)

    walk(0, 0, 1 << (max(w, h) - 1).bit_length())  # the quadtree root: next power of two
    return out


def _exhaustive(ctx):
    return {(s, sh): {(r, c): x for r, c, x in zip(rows, cols, cands)}
            for s, sh, rows, cols, cands, _ in ctx.tables()}


@pytest.mark.parametrize("w,h,seed,mn,mx,shifts,t_rms", [
    (160, 96, 5, 4, 16, [4], 8.0),
    (75, 53, 6, 4, 32, [4], 8.0),        # odd sides: border splits down to size 2
    (128, 128, 7, 4, 32, [4, 8], 3.0),   # a second stride is never asked
    (96, 64, 8, 2, 16, [4], 20.0),
])
def test_walk_tables_hold_exactly_the_walk(w, h, seed, mn, mx, shifts, t_rms):
    img = synth.plane(w, h, seed)
    p = EncodeParams(min_size=mn, max_size=mx, shift=shifts[0])
    print(f"walk tables {w}x{h} sizes {mn}..{mx} shifts {shifts} t_rms {t_rms}", flush=True)
    full = HipContext(0)
    full.prepare(img, p, shifts, method=FISHER)
    want = _exhaustive(full)
    fis_want = full.method_tables(FISHER)
    vis = _visited(w, h, p, shifts[0], t_rms, want)
    assert vis
    ctx = HipContext(0)
    ctx.set_walk_tables(t_rms)
    ctx.prepare(img, p, shifts, method=FISHER)
    for it in range(2):  # the prepare's step, then graph replays
        n_vis = 0
        for (size, sh), table in want.items():
            for (row, col), cand in table.items():
                rc, got, ev = ctx.search_method(EXH, row, col, size, sh)
                if sh == shifts[0] and (size, row, col) in vis:
                    n_vis += 1
                    assert rc == N.OK, (size, sh, row, col)
                    assert _key(got) == _key(cand), (size, sh, row, col)
                else:
                    assert rc == N.E_NOT_IN_TABLE, (size, sh, row, col, rc)
        assert n_vis == len(vis)
        assert {k: [_key(c) for c in v[2]] for k, v in ctx.method_tables(FISHER).items()} == \
            {k: [_key(c) for c in v[2]] for k, v in fis_want.items()}
        ctx.run(iters=3, use_graph=True)
    ctx.close()
    full.close()


def test_walk_tables_need_fisher_and_winners_only():
    img = synth.plane(96, 64, 9)
    p = EncodeParams(min_size=4, max_size=16, shift=4)
    ctx = HipContext(0)
    ctx.set_walk_tables(8.0)
    ctx.prepare(img, p, [4])  # exhaustive method: full tables
    assert all(ctx.search_method(EXH, r, c, s, 4)[0] == N.OK
               for s, sh, rows, cols, _, _ in ctx.tables() for r, c in zip(rows, cols))
    ctx.set_ranked(3)
    ctx.prepare(img, p, [4], method=FISHER)  # ranked: full tables
    assert all(ctx.search_method(EXH, r, c, s, 4)[0] == N.OK
               for s, sh, rows, cols, _, _ in ctx.tables() for r, c in zip(rows, cols))
    ctx.set_ranked(1)
    ctx.set_walk_tables(0.0)
    ctx.prepare(img, p, [4], method=FISHER)
    assert all(ctx.search_method(EXH, r, c, s, 4)[0] == N.OK
               for s, sh, rows, cols, _, _ in ctx.tables() for r, c in zip(rows, cols))
    ctx.close()
