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
