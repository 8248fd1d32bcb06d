"""The MFMA search's stage-1 limit, restated in binary32 on the CPU.

k_search_mfma drops a candidate in stage 1 when
    t1_x4 < floor((S1*t0 + nt) / n),
nt a binary32 lower bound of the smallest Ni = n*t1 - S1*t0 at which the
candidate's quantised SSE could still reach the threshold tau (encode.rs:247-289
fit_f64 gives the SSE).  Sizes <= 4 add the slope cap: alfa2 of
encode.rs:266 is at most qa_max * max_alfa / 2^bits_alfa, which raises the
bound to the tangent (A - n tau + U^2 Di) / (2U) where A - n tau > U^2 Di.

The check: every candidate of a small plane must pass the limit computed
from its own inflated SSE (the tightest threshold it can ever meet), with the
kernel's operations and roundings.  Candidates whose quantised slope sits at
the cap come within a few units of the limit, so the same check with a cap
10% too small (an over-pruning bound) must fail.

Sizes 4 and 8 test stage 1 on the matrix pipe instead: the range operand is
centred (r - m, m = round(t0/n)), so the correlation K-steps give the exact
X0 = t1 - m S1, and one more binary16 K-step gives, per range and position,
four right-hand sides nb_j = S1 e/n + T_j(sqrt(Di))/n - M_j (e = t0 - n m):
T_0 the plain bound x*sqrt(Di) and T_1..3 tangents of the capped bound at
sqrt(Di) = 0, x/(3U), 2x/(3U), every factor rounded down into binary16
pieces, M_j a rounding margin; a candidate passes when X0 > max_j nb_j.  The
same own-threshold check applies to each nb_j: X0 must exceed its exact value
by at least M_j, and M_j must exceed the worst rounding of the K-step; the
slope and d taken as single binary16 values rounded to nearest, or a cap 10%
too small in the tangents, must be caught.

k_fisher_search skips a bucket candidate's binary64 fit when the free
least-squares bound, A Di - max(Ni, 0)^2 > n tau Di (+ a 1e-12 A Di margin),
rules it out against the wave's best inflated SSE; the same own-threshold
check applies, and dropping the n factor must be caught.
"""

import numpy as np
import pytest
from numpy.lib.stride_tricks import sliding_window_view as swv

from mars_amd import synth
from tests import _oracle as O

f32 = np.float32


def kernel_constants(max_alfa, bits_alfa):
    """cap_u and cap_k as the context computes them (rounded up / down)."""
    qa_max = (1 << bits_alfa) - 1
    u = qa_max * (max_alfa / 2.0 ** bits_alfa) * 0.25 * (1.0 + 1e-12)
    cap_u = np.nextafter(f32(u), f32(np.inf))
    k = 1.0 / (2.0 * float(cap_u)) if u > 0 else 8.0
    cap_k = np.nextafter(f32(min(k, 8.0)), f32(0.0))
    return cap_u, cap_k


def tau_f32(rms, n):
    """tau_bits of the candidate's SSE (sum = n rms^2 up to binary64 rounding)."""
    s = rms * rms * n
    return f32(s * (1.0 + 1e-9) + 1e-6 * n) * (1.0 + 1e-7)


def limit(A, Di, S1, t0, thr, n, cap_u, cap_k, use_cap):
    """floor((S1*t0 + nt) / n) with the kernel's binary32 steps; None: no pruning."""
    alo = f32(f32(A) * f32(1.0 - 3e-7))
    amt = f32(alo - f32(n) * f32(thr))
    if not amt > 0:
        return None
    sa = f32(f32(np.sqrt(amt)) * f32(1.0 - 2e-5))
    sq = f32(np.sqrt(float(Di)))  # sqrt(Di) rounded down to binary32
    if float(sq) > np.sqrt(float(Di)):
        sq = np.nextafter(sq, f32(0.0))
    nt = f32(sa * sq)
    if use_cap:
        d = max(f32(sa - f32(cap_u * sq)), f32(0.0))
        dd = f32(d * d)
        nt = f32(float(dd) * float(cap_k) + float(nt))
    return (int(S1) * int(t0) + int(nt)) // n


def candidates(size, seed):
    img = synth.plane(48, 40, seed)
    d = O.contract(img).astype(np.int64)
    n = size * size
    W = swv(d, (size, size))[::2, ::2]
    wins = W.reshape(-1, n)
    S1 = wins.sum(1)
    S2 = (wins * wins).sum(1)
    for r0 in range(0, 40 - size + 1, 2 * size):
        for c0 in range(0, 48 - size + 1, 3 * size):
            rng = img[r0:r0 + size, c0:c0 + size]
            rr = rng.astype(np.int64).reshape(n)
            t0, t2 = int(rr.sum()), int((rr * rr).sum())
            for k in range(8):
                rk = O.permute_range(rng, k).astype(np.int64).reshape(n)
                t1 = wins @ rk
                for p in range(len(wins)):
                    yield (n, int(S1[p]), int(S2[p]), t0, int(t1[p]), t2)


@pytest.mark.parametrize("size,seed,max_alfa,bits_alfa", [
    (4, 3, 1.0, 4), (4, 4, 1.0, 4), (2, 5, 1.0, 4), (4, 6, 0.5, 3), (4, 7, 0.75, 5),
])
def test_every_candidate_passes_its_own_threshold(size, seed, max_alfa, bits_alfa):
    cap_u, cap_k = kernel_constants(max_alfa, bits_alfa)
    bad_u = f32(cap_u * 0.9)
    bad_k = f32(1.0 / (2.0 * float(bad_u)))
    n_cand = n_capped = misses_bad = tight = 0
    for m in candidates(size, seed):
        n, S1, S2, t0, t1, t2 = m
        qa, _, rms = O.fit_f64(m, max_alfa, bits_alfa, 7)
        thr = tau_f32(rms, n)
        A, Di = n * t2 - t0 * t0, n * S2 - S1 * S1
        n_cand += 1
        lim_plain = limit(A, Di, S1, t0, thr, n, cap_u, cap_k, False)
        lim = limit(A, Di, S1, t0, thr, n, cap_u, cap_k, True)
        if lim is None:
            continue
        assert t1 >= lim >= lim_plain, (m, qa, rms, lim, lim_plain)
        if lim > lim_plain:
            n_capped += 1
            tight += t1 - lim < 16
        if t1 < limit(A, Di, S1, t0, thr, n, bad_u, bad_k, True):
            misses_bad += 1
    print(f"{n_cand} candidates, {n_capped} with a raised limit, {tight} within 16 of it, "
          f"{misses_bad} dropped by the 10% smaller cap")
    assert n_cand > 10000 and n_capped > 100, (n_cand, n_capped)
    assert tight > 0, "no candidate came near the capped limit"
    assert misses_bad > 0, "an over-pruning cap went unnoticed"


def fisher_skips(m, rms, n_factor=True):
    """k_fisher_search's test: skip the fit when A Di - Ni+^2 > n tau Di + 1e-12 A Di."""
    n, S1, S2, t0, t1, t2 = m
    tau = max(rms * rms * n, 0.0) * (1.0 + 1e-9) + 1e-6 * n
    A = n * t2 - t0 * t0
    Ni, Di = n * t1 - S1 * t0, n * S2 - S1 * S1
    if Di <= 0:
        return False
    ad = float(A) * float(Di)
    np_ = float(Ni) if Ni > 0 else 0.0
    return ad - np_ * np_ > (float(n) if n_factor else 1.0) * tau * float(Di) + 1e-12 * ad


@pytest.mark.parametrize("size,seed", [(4, 3), (8, 4), (2, 5), (16, 6)])
def test_fisher_bound_never_skips_a_candidate_at_its_own_threshold(size, seed):
    n_cand = caught = 0
    for m in candidates(size, seed):
        _, _, rms = O.fit_f64(m, 1.0, 4, 7)
        assert not fisher_skips(m, rms), m
        caught += fisher_skips(m, rms, n_factor=False)
        n_cand += 1
    assert n_cand >= 100
    assert caught > 0, "a bound without the n factor went unnoticed"


f16 = np.float16


def f16_rd(x):
    """The largest binary16 <= x (x >= 0 a binary32)."""
    h = f16(x)
    if float(h) > float(x):
        h = np.nextafter(h, f16(-np.inf))
    return h


def f16_ru(x):
    h = f16(x)
    if float(h) < float(x):
        h = np.nextafter(h, f16(np.inf))
    return h


def bias_gaps(m, thr, cap_k, split=True):
    """k_search_mfma's binary16 stage-1 right-hand sides for one candidate:
    per bound j, (X0 minus the exact nb_j, M_j, the sum of the K-step's term
    magnitudes); None when the range cannot prune.  split=False: the slope
    and d as one binary16 each, rounded to nearest (an over-pruning variant)."""
    import math
    n, S1, S2, t0, t1, t2 = m
    lg = n.bit_length() - 1
    A, Di = n * t2 - t0 * t0, n * S2 - S1 * S1
    mc = (t0 + n // 2) >> lg
    e = t0 - n * mc
    x0 = t1 - mc * S1
    # position record (k_possums)
    sq = f32(np.sqrt(float(Di)))
    if float(sq) > np.sqrt(float(Di)):
        sq = np.nextafter(sq, f32(0.0))
    dn = f32(f32(sq * f32(1.0 - 2.0 ** -20)) * f32(1.0 / n))
    dh = f16_rd(dn) if split else f16(dn)
    dl = f16_rd(f32(dn - f32(dh))) if split else f16(0)
    b = [(S1 >> 8) * 256 / n, (S1 & 255) / n, float(dh), float(dl), 2048.0, 1.0, 1.0, 1.0, float(dh)]
    assert all(float(f16(v)) == v for v in b[:2])
    # range operand (make_abias), one row per bound
    alo = f32(f32(A) * f32(1.0 - 3e-7))
    amt = f32(alo - f32(n) * f32(thr))
    if not amt > 0:
        return None
    x = f32(np.sqrt(amt))
    be0 = f32(1020.0 * abs(e))
    rd = f16_rd if split else f16  # toward zero (v_cvt_pkrtz) = down for these values
    out = []
    for tf in (f32(0.0), f32(1.0), f32(2.0 / 3.0), f32(1.0 / 3.0)):
        bsl = f32(f32(f32(1.0) - tf) * f32(1.0 - 2.0 ** -19))
        bkn = f32(f32(f32(f32(tf * f32(f32(2.0) - tf)) * f32(cap_k)) * f32(1.0 / n)) * f32(1.0 - 2.0 ** -19))
        sg = f32(x * bsl)
        kn = f32(amt * bkn)
        sh = rd(sg)
        k1 = f32(f32(f16_rd(f32(kn * f32(2.0 ** -11)))) * f32(2048.0))
        r1 = f32(kn - k1)
        assert float(r1) == float(kn) - float(k1)
        sl = f16_rd(f32(sg - f32(sh))) if split else f16(0)
        k2 = f16_rd(r1)
        r2 = f32(r1 - f32(k2))
        k3 = f16_rd(r2)
        # the margin of the largest x (tau = 0), per range
        sa = f32(f32(np.sqrt(max(alo, f32(0.0)))) * f32(1.0 + 2.0 ** -10))
        mm = f32(f32(f32(be0 + f32(f32(f32(1020.0) * sa) * bsl)) + f32(f32(alo * f32(1.0 + 2.0 ** -10)) * bkn))
                 * f32(2.0 ** -19))
        mm = max(f32(mm * f32(1.0 + 2.0 ** -8)), f32(2.0 ** -14))
        M = float(f16_rd(mm))
        a = [e, e, float(sh), float(sh), float(k1) / 2048.0, float(k2), float(k3), -M, float(sl)]
        assert all(float(f16(v)) == v for v in a)
        terms = [float(ai) * bi for ai, bi in zip(a, b)]
        out.append((x0 - math.fsum(terms), M, sum(abs(v) for v in terms)))
    return out


@pytest.mark.parametrize("size,seed", [(4, 3), (4, 8), (8, 4), (8, 9)])
def test_binary16_bias_keeps_every_candidate_at_its_own_threshold(size, seed):
    cap_u, cap_k = kernel_constants(1.0, 4)
    bad_k = f32(1.0 / (2.0 * float(f32(cap_u * 0.9))))
    n_cand = tight = caught = caught_cap = raised = 0
    for m in candidates(size, seed):
        n = m[0]
        _, _, rms = O.fit_f64(m, 1.0, 4, 7)
        thr = tau_f32(rms, n)
        r = bias_gaps(m, thr, cap_k)
        if r is None:
            continue
        n_cand += 1
        for gap, M, mag in r:
            # the exact pieces never exceed the bound: X0 - nb_j >= M_j
            assert gap >= M, (m, gap, M)
            # M_j exceeds 8 roundings of an ulp of the largest partial sum
            ulp = 2.0 ** (np.floor(np.log2(mag)) - 23) if mag > 0 else 0.0
            assert M > 8 * ulp, (m, gap, M, mag)
        best = min(g - M for g, M, _ in r)
        tight += best < 16.0
        raised += min(g - M for g, M, _ in r[1:]) < r[0][0] - r[0][1]
        wrong = bias_gaps(m, thr, cap_k, split=False)
        caught += any(g < M for g, M, _ in wrong)
        wrong_cap = bias_gaps(m, thr, bad_k)
        caught_cap += any(g < M for g, M, _ in wrong_cap)
    print(f"{n_cand} candidates, {raised} raised by a tangent, {tight} within 16 of a bound, "
          f"{caught} lost with unsplit factors, {caught_cap} lost with a 10% smaller cap")
    assert n_cand > 1000
    assert tight > 0, "no candidate came near the bounds"
    assert caught > 0, "single binary16 slope and d went unnoticed"
    if size == 4:  # at size 8 the candidates of these planes sit beyond the cap's reach
        assert raised > 0, "no tangent ever raised the bound"
        assert caught_cap > 0, "an over-pruning cap in the tangents went unnoticed"
