"""Pins of the CPU oracle (oracle/search.c) -- CPU only.

What fixes each oracle function, independently of its own code:
  * fit_f64          flat128 known answer (crates/mars-codec/src/encode.rs:2207-2231),
                     best_beta worked example (2233-2241), and exact rational
                     arithmetic (fractions) for qalfa / qbeta / SSE on random
                     moments -- a dropped term, a wrong sign or swapped s1/t0
                     fails those.
  * fit_f32          exact rational arithmetic with per-operation binary32
                     error bounds at s <= 4 (codes away from rounding
                     boundaries, SSE of the returned codes); planted mistakes
                     (a dropped SSE term, s1*t1 for s1*t0) fail it.
  * contract / domain_sums / cross_term / permute_range
                     the reference's own exactness construction
                     (encode.rs:2101-2205: integer moments == the mean-based
                     f64 definition), restated on random blocks, and numpy's
                     rot90 / flip / transpose for the 8 isometries.
  * search_top_k     brute force over every (position, isometry) on small
                     planes, evals == positions * 8, strict-< first-minimum
                     tie rule and ascending top-k (encode.rs:2988-3032).
"""

from fractions import Fraction

import numpy as np
import pytest

from tests import _oracle as O


# ---------------------------------------------------------------- known answers

def test_flat128_fit_known_answer():
    size = 16
    n = size * size
    m = (n, 128 * 4 * n, 512 * 512 * n, 128 * n, 128 * 512 * n, 128 * 128 * n)
    qa, qb, rms = O.fit_f64(m, 1.0, 4, 7)
    assert qa == 0
    assert qb == 64
    assert abs(rms - 0.5039370078736174) < 1e-12


def test_best_beta_worked_example():
    assert O.lib().or_best_beta(128 * 256, 256, 7) == 64


def test_fit_f32_flat128():
    size = 16
    n = size * size
    m = (n, 128 * 4 * n, 512 * 512 * n, 128 * n, 128 * 512 * n, 128 * 128 * n)
    qa, qb, rms = O.fit_f32(m, 1.0, 4, 7)
    assert (qa, qb) == (0, 64)
    # binary32 cancellation in `sum` costs about 1e-4 here (the D23 effect).
    assert abs(rms - 0.50393701) < 1e-3


# ---------------------------------------------------------------- exact-arithmetic pins

def _round_half_up(x: Fraction) -> int:
    # trunc(0.5 + x) for x >= 0
    return int((x + Fraction(1, 2)).__floor__())


def _exact_fit(m, max_alfa=1, bits_alfa=4, bits_beta=7):
    """Exact-rational statement of the quantised fit (what binary64 approximates)."""
    s0, s1x4, s2x16, t0, t1x4, t2 = m
    s1, s2, t1 = Fraction(s1x4, 4), Fraction(s2x16, 16), Fraction(t1x4, 4)
    det = s0 * s2 - s1 * s1
    alfa = Fraction(0) if det == 0 else (s0 * t1 - s1 * t0) / det
    alfa = max(alfa, Fraction(0))
    qmax = (1 << bits_alfa) - 1
    qa = min(_round_half_up(alfa / max_alfa * (1 << bits_alfa)), qmax)
    a2 = Fraction(qa, 1 << bits_alfa) * max_alfa
    beta = (t0 - a2 * s1) / s0
    if a2 > 0:
        beta += a2 * 255
    bmax = (1 << bits_beta) - 1
    y = beta / ((1 + abs(a2)) * 255) * bmax
    qb = min(max(_round_half_up(y), 0), bmax)
    b2 = Fraction(qb, bmax) * ((1 + abs(a2)) * 255)
    if a2 > 0:
        b2 -= a2 * 255
    sse = t2 - 2 * a2 * t1 - 2 * b2 * t0 + a2 * a2 * s2 + 2 * a2 * b2 * s1 + s0 * b2 * b2
    return qa, qb, sse, alfa * (1 << bits_alfa), y


def _pixel_sse(rng, dom, qa, qb, bits_alfa=4, bits_beta=7):
    """SSE of the quantised affine map, pixel by pixel, in exact rationals."""
    a2 = Fraction(qa, 1 << bits_alfa)
    bmax = (1 << bits_beta) - 1
    b2 = Fraction(qb, bmax) * ((1 + a2) * 255)
    if a2 > 0:
        b2 -= a2 * 255
    tot = Fraction(0)
    for r, d4 in zip(rng.ravel().tolist(), dom.ravel().tolist()):
        e = r - a2 * Fraction(d4, 4) - b2
        tot += e * e
    return tot


@pytest.mark.parametrize("size", [2, 4, 8])
def test_fit_f64_against_exact_rationals(size):
    rng = np.random.default_rng(100 + size)
    checked = 0
    for trial in range(300):
        # Mix textures: random, affine-related (large alfa), flat.
        rr = rng.integers(0, 256, size=(size, size))
        dd = rng.integers(0, 1021, size=(size, size))
        if trial % 3 == 1:
            a = rng.integers(1, 16) / 16.0
            rr = np.clip(np.rint(a * dd / 4.0 + rng.integers(-20, 60)
                                 + rng.normal(0, 2, size=(size, size))), 0, 255).astype(int)
        elif trial % 3 == 2:
            rr = np.full((size, size), int(rng.integers(0, 256)))
        n = size * size
        m = (n, int(dd.sum()), int((dd.astype(np.int64) ** 2).sum()), int(rr.sum()),
             int((rr.astype(np.int64) * dd).sum()), int((rr.astype(np.int64) ** 2).sum()))
        qa, qb, rms = O.fit_f64(m)
        eqa, eqb, sse, x, y = _exact_fit(m)
        # Quantisation decisions are exact unless the rational sits on a half-integer.
        if (x + Fraction(1, 2)).denominator != 1:
            assert qa == eqa, (m, qa, eqa)
        if qa == eqa and (y + Fraction(1, 2)).denominator != 1:
            assert qb == eqb, (m, qb, eqb)
        # The binary64 SSE agrees with the pixel-wise rational SSE of the same codes.
        psse = _pixel_sse(rr, dd, qa, qb)
        assert abs(rms * rms * n - float(psse)) <= 1e-7 * max(1.0, float(psse)) + 1e-6
        checked += 1
    assert checked == 300


@pytest.mark.parametrize("size", [2, 4])
def test_fit_f32_against_exact_rationals(size):
    """or_fit_f32 (encode.rs:295-337, the FP32 baseline's fit) against exact
    rationals.  At s <= 4 every moment is an integer below 2^24, exact in
    binary32, so the only errors are the fit's own roundings; each is bounded
    here from the operands (unit roundoff 2^-24 per operation, with margin).
    Codes must be the exact ones wherever the exact rational lies farther
    from a rounding boundary than that bound, and s0 * rms^2 must be the exact
    SSE of the returned codes within the bound of the `sum` expression -- a
    wrong alfa branch, a swapped operand or a dropped term is off by orders of
    magnitude more."""
    u = 2.0 ** -24
    rng = np.random.default_rng(200 + size)
    n = size * size
    codes_checked = 0
    for trial in range(600):
        rr = rng.integers(0, 256, size=(size, size))
        dd = rng.integers(0, 1021, size=(size, size))
        if trial % 3 == 1:
            a = rng.integers(1, 16) / 16.0
            rr = np.clip(np.rint(a * dd / 4.0 + rng.integers(-20, 60)
                                 + rng.normal(0, 2, size=(size, size))), 0, 255).astype(int)
        elif trial % 3 == 2:
            rr = np.full((size, size), int(rng.integers(0, 256)))
        m = (n, int(dd.sum()), int((dd.astype(np.int64) ** 2).sum()), int(rr.sum()),
             int((rr.astype(np.int64) * dd).sum()), int((rr.astype(np.int64) ** 2).sum()))
        assert max(m) < 2 ** 24
        qa, qb, rms = O.fit_f32(m)
        eqa, eqb, _, x, y = _exact_fit(m)
        s0, s1, s2, t0, t1, t2 = (m[0], m[1] / 4, m[2] / 16, m[3], m[4] / 4, m[5])
        num, den = s0 * t1 - s1 * t0, s0 * s2 - s1 * s1
        if den != 0 and num > 0:
            ex = abs(float(x)) * 8 * u * ((abs(s0 * t1) + abs(s1 * t0)) / abs(num)
                                          + (s0 * s2 + s1 * s1) / abs(den) + 2)
            if abs(float(x) - (np.floor(float(x)) + 0.5)) > 8 * ex + 1e-9:
                assert qa == eqa, (m, qa, eqa)
                codes_checked += 1
        elif num <= 0 or den == 0:
            assert qa == 0, m
        if qa == eqa:
            a2 = qa / 16.0
            eb = 8 * u * ((abs(t0) + abs(a2 * s1)) / s0 + a2 * 255 + abs(float(y)))
            ey = eb / ((1 + a2) * 255) * 127 + 8 * u * abs(float(y))
            if abs(float(y) - (np.floor(float(y)) + 0.5)) > 8 * ey + 1e-9:
                assert qb == eqb, (m, qb, eqb)
        a2 = Fraction(qa, 16)
        b2 = Fraction(qb, 127) * ((1 + a2) * 255) - (a2 * 255 if a2 > 0 else 0)
        terms = [t2, 2 * a2 * Fraction(m[4], 4), 2 * b2 * t0, a2 * a2 * Fraction(m[2], 16),
                 2 * a2 * b2 * Fraction(m[1], 4), s0 * b2 * b2]
        sse = terms[0] - terms[1] - terms[2] + terms[3] + terms[4] + terms[5]
        bound = 32 * u * sum(abs(float(t)) for t in terms) + 4 * u * s0 * rms * rms + 1e-6
        assert abs(s0 * rms * rms - float(sse)) <= bound, (m, rms, float(sse) / s0)
    assert codes_checked > 250


def test_fit_f64_perfect_affine_block():
    # r = (q/16) * d exactly (d a multiple of 4*16): alfa quantises to q.
    size = 4
    n = size * size
    rng = np.random.default_rng(7)
    for q in range(1, 16):
        d4 = rng.integers(0, 16, size=(size, size)) * 64  # D values, d = D/4 multiple of 16
        r = (q * d4) // 64
        m = (n, int(d4.sum()), int((d4 ** 2).sum()), int(r.sum()), int((r * d4).sum()),
             int((r ** 2).sum()))
        qa, qb, rms = O.fit_f64(m)
        assert qa == q


# ---------------------------------------------------------------- moments

def _mean_moments(rng_blk, dom_px, size, k):
    """encode.rs:2142-2177: the real-valued definition with d = mean of 4 pixels."""
    d = np.zeros((size, size))
    for u in range(size):
        for v in range(size):
            d[u, v] = (float(dom_px[2 * u, 2 * v]) + float(dom_px[2 * u, 2 * v + 1]) +
                       float(dom_px[2 * u + 1, 2 * v]) + float(dom_px[2 * u + 1, 2 * v + 1])) / 4.0
    rk = O.permute_range(rng_blk, k).astype(np.float64)
    return float(d.sum()), float((d * d).sum()), float(rk.sum()), float((rk * d).sum()), \
        float((rk * rk).sum())


def test_integer_moments_equal_the_mean_definition():
    gen = np.random.default_rng(0x5EED)
    for _ in range(400):
        size = int(gen.choice([2, 4, 8, 16]))
        img = gen.integers(0, 256, size=(4 * size, 4 * size), dtype=np.uint8)
        row, col = size * int(gen.integers(0, 4)), size * int(gen.integers(0, 4))
        dom_row, dom_col = 2 * int(gen.integers(0, size + 1)), 2 * int(gen.integers(0, size + 1))
        k = int(gen.integers(0, 8))
        s0, s1, s2, t0, t1, t2 = O.integer_moments(img, row, col, size, dom_row, dom_col, k)
        ms1, ms2, mt0, mt1, mt2 = _mean_moments(
            img[row:row + size, col:col + size],
            img[dom_row:dom_row + 2 * size, dom_col:dom_col + 2 * size], size, k)
        assert s0 == size * size
        assert s1 / 4.0 == ms1 and s2 / 16.0 == ms2
        assert t0 == mt0 and t1 / 4.0 == mt1 and t2 == mt2


def test_isometries_match_numpy_symmetries():
    x = np.arange(36, dtype=np.uint8).reshape(6, 6)
    expect = {
        0: x,
        1: np.rot90(x, -1),
        2: np.rot90(x, 1),
        3: np.rot90(x, 2),
        4: np.fliplr(x),
        5: np.flipud(x),
        6: x.T,
        7: np.rot90(x, 2).T,
    }
    for k, e in expect.items():
        assert np.array_equal(O.permute_range(x, k), e), k
    perms = {O.permute_range(x, k).tobytes() for k in range(8)}
    assert len(perms) == 8


def test_contract_brute_force_odd_sizes():
    img = np.random.default_rng(3).integers(0, 256, size=(7, 9), dtype=np.uint8)
    d = O.contract(img)
    assert d.shape == (3, 4)
    for i in range(3):
        for j in range(4):
            assert d[i, j] == int(img[2 * i:2 * i + 2, 2 * j:2 * j + 2].astype(int).sum())


# ---------------------------------------------------------------- search

def _brute(img, row, col, size, shift):
    h, w = img.shape
    best = None
    evals = 0
    order = 0
    for dom_row in range(0, h - 2 * size + 1, shift):
        for dom_col in range(0, w - 2 * size + 1, shift):
            for k in range(8):
                m = O.integer_moments(img, row, col, size, dom_row, dom_col, k)
                qa, qb, rms = O.fit_f64(m)
                evals += 1
                if best is None or rms < best[0]:
                    best = (rms, dom_row, dom_col, k, qa, qb, m)
                order += 1
    return best, evals


@pytest.mark.parametrize("shift", [2, 3, 4])
def test_search_top_k_matches_brute_force(shift):
    img = np.random.default_rng(11 + shift).integers(0, 256, size=(20, 18), dtype=np.uint8)
    for (row, col, size) in [(0, 0, 4), (8, 4, 4), (4, 8, 2), (0, 8, 8)]:
        best, evals = _brute(img, row, col, size, shift)
        got, ev = O.search_top_k(img, row, col, size, shift)
        assert ev == evals
        c = got[0]
        assert (c.rms, c.dom_row, c.dom_col, c.isometry, c.qalfa, c.qbeta) == best[:6]
        assert (c.m.s0, c.m.s1_x4, c.m.s2_x16, c.m.t0, c.m.t1_x4, c.m.t2) == best[6]


def test_first_minimum_wins_ties_on_a_flat_plane():
    img = np.full((32, 32), 77, dtype=np.uint8)
    got, ev = O.search_top_k(img, 8, 8, 4, 4)
    c = got[0]
    # Every candidate ties (det == 0): the first enumerated one wins.
    assert (c.dom_row, c.dom_col, c.isometry) == (0, 0, 0)
    assert ev == ((32 - 8) // 4 + 1) ** 2 * 8


def test_top_k_ascending_and_rank0_is_winner():
    img = np.random.default_rng(5).integers(0, 256, size=(64, 64), dtype=np.uint8)
    top1, ev1 = O.search_top_k(img, 0, 0, 16, 4, k=1)
    top3, ev3 = O.search_top_k(img, 0, 0, 16, 4, k=3)
    assert ev1 == ev3 == ((64 - 32) // 4 + 1) ** 2 * 8
    assert top3[0].rms == top1[0].rms
    assert (top3[0].dom_row, top3[0].dom_col, top3[0].isometry) == \
        (top1[0].dom_row, top1[0].dom_col, top1[0].isometry)
    assert top3[0].rms <= top3[1].rms <= top3[2].rms
    assert len({(c.dom_row, c.dom_col, c.isometry) for c in top3}) == 3


def test_no_legal_domain():
    img = np.zeros((8, 8), dtype=np.uint8)
    got, ev = O.search_top_k(img, 0, 0, 8, 4)
    assert got == [] and ev == 0


def test_table_is_thread_count_invariant():
    img = np.random.default_rng(9).integers(0, 256, size=(64, 64), dtype=np.uint8)
    nodes = O.nodes(64, 64, 4, 16)
    rows = [r for r, _ in nodes[8]]
    cols = [c for _, c in nodes[8]]
    a, ea = O.search_table(img, 8, 4, rows, cols, threads=1)
    b, eb = O.search_table(img, 8, 4, rows, cols, threads=5)
    assert ea == eb
    for x, y in zip(a, b):
        assert (x.rms, x.dom_row, x.dom_col, x.isometry) == (y.rms, y.dom_row, y.dom_col,
                                                             y.isometry)


def test_winner_consistency_check_pins_moments_and_fit():
    """The table check used at sizes the oracle cannot search in full
    (tests/_checks.py) accepts the oracle's own winners and rejects a winner
    whose isometry, position or rms was altered."""
    from mars_amd import synth
    from tests._checks import winner_consistency

    img = synth.plane(64, 48, 9)
    rows = [r for r in range(0, 48, 8) for _ in range(0, 64, 8)]
    cols = [c for _ in range(0, 48, 8) for c in range(0, 64, 8)]
    ref, _ = O.search_table(img, 8, 2, rows, cols, threads=4)
    assert winner_consistency(img, 8, rows, cols, ref) == len(rows)
    for field, delta in (("isometry", 1), ("dom_col", 2), ("rms", 1e-12)):
        bad = list(ref)
        c = O.OrCandidate.from_buffer_copy(ref[5])
        v = getattr(c, field)
        setattr(c, field, (v + delta) % 8 if field == "isometry" else v + delta)
        bad[5] = c
        with pytest.raises(AssertionError):
            winner_consistency(img, 8, rows, cols, bad)
