"""Size-independent checks of search tables (test infrastructure)."""

import struct

import numpy as np

from tests import _oracle as O


def _bits(x: float) -> int:
    return struct.unpack("<q", struct.pack("<d", x))[0]


def contract(img: np.ndarray) -> np.ndarray:
    """encode.rs:467-528 Contracted: 2x2 box sums, odd last row/column dropped."""
    h, w = img.shape
    x = img[: 2 * (h // 2), : 2 * (w // 2)].astype(np.int64)
    return x[0::2, 0::2] + x[1::2, 0::2] + x[0::2, 1::2] + x[1::2, 1::2]


def winner_consistency(img, size, rows, cols, cands):
    """Checks of every winner of one table that hold at any plane size: a legal
    domain position and isometry, RawMoments equal to the window and range sums
    computed here from the pixels (numpy box sums; t1 through the oracle's
    permute_range), and codes and rms equal to the oracle's fit_f64 of those
    moments (encode.rs:247-289).  Says nothing about whether the winner is the
    minimum: sampled oracle searches check that."""
    h, w = img.shape
    d = contract(img)
    n = len(rows)
    dr = np.array([c.dom_row for c in cands], np.int64)
    dc = np.array([c.dom_col for c in cands], np.int64)
    assert dr.max() <= h - 2 * size and dc.max() <= w - 2 * size, "winner outside the pool"
    assert max(c.isometry for c in cands) < 8
    ii = np.arange(size)
    win = d[(dr // 2)[:, None, None] + ii[None, :, None],
            (dc // 2)[:, None, None] + ii[None, None, :]]
    r0 = np.array(rows, np.int64)
    c0 = np.array(cols, np.int64)
    rng = img[r0[:, None, None] + ii[None, :, None], c0[:, None, None] + ii[None, None, :]]
    perm = np.stack([O.permute_range(rng[k], cands[k].isometry) for k in range(n)])
    perm = perm.astype(np.int64)
    r64 = rng.astype(np.int64)
    want = np.stack([np.full(n, size * size), win.sum((1, 2)), (win * win).sum((1, 2)),
                     r64.sum((1, 2)), (perm * win).sum((1, 2)), (r64 * r64).sum((1, 2))], 1)
    ms = [c.m if hasattr(c, "m") else c for c in cands]
    got = np.array([(m.s0, m.s1_x4, m.s2_x16, m.t0, m.t1_x4, m.t2) for m in ms], np.int64)
    bad = np.nonzero((got != want).any(1))[0]
    assert bad.size == 0, f"{bad.size} winners' moments differ, first range {bad[0]}"
    for k, c in enumerate(cands):
        qa, qb, rms = O.fit_f64(tuple(int(x) for x in want[k]))
        assert (c.qalfa, c.qbeta, _bits(c.rms)) == (qa, qb, _bits(rms)), f"range {k} fit"
    return n
