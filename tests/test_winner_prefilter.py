"""The search drains' per-range winner pre-filter (mars_amd/csrc/kernels.hip,
process()): of a drain's winners, only those whose binary64 rms has the
least high 32-bit word in their range stay winners before the serial
best-record loop.  That keeps each range's (rms, key) minimum because a
non-negative binary64 orders as its bit pattern, so its high word is
monotone in the value; checked here on random and adversarial rms sets
(equal values, equal high words with different low words, +inf, 0)."""

import numpy as np


def high_word(x):
    return int(np.array([x], dtype=np.float64).view(np.uint64)[0] >> np.uint64(32))


def prefilter_then_serial(winners):
    """winners: list of (slot, rms, key) -> {slot: (rms, key)} as the kernel resolves them."""
    hmin = {}
    for slot, rms, key in winners:
        hw = high_word(rms)
        hmin[slot] = min(hmin.get(slot, 0xFFFFFFFF), hw)
    kept = [(s, r, k) for s, r, k in winners if high_word(r) == hmin[s]]
    best = {}
    for s, r, k in kept:  # the serial loop: strict (rms, key) order
        if s not in best or (r, k) < best[s]:
            best[s] = (r, k)
    return best, len(kept)


def test_prefilter_keeps_every_ranges_minimum():
    rng = np.random.default_rng(5)
    base = [0.0, 1.0, 1.0 + 2.0 ** -40, 1.0 + 2.0 ** -20, 3.5, 3.5, np.inf, 1e-300, 7.25]
    dropped = 0
    for trial in range(3000):
        n = int(rng.integers(1, 64))
        winners = []
        for lane in range(n):
            slot = int(rng.integers(0, 8))
            if rng.random() < 0.5:
                rms = float(rng.choice(base))
            else:  # same high word, different low words
                rms = float(np.array([0x4010000000000000 + int(rng.integers(0, 1 << 20))],
                                     dtype=np.uint64).view(np.float64)[0])
            winners.append((slot, rms, int(rng.integers(0, 1 << 20))))
        got, kept = prefilter_then_serial(winners)
        want = {}
        for s, r, k in winners:
            if s not in want or (r, k) < want[s]:
                want[s] = (r, k)
        assert got == want, (trial, winners)
        dropped += n - kept
    assert dropped > 0, "the pre-filter never removed a winner: the test does not exercise it"


def test_high_word_is_monotone_on_non_negative_doubles():
    rng = np.random.default_rng(6)
    xs = np.sort(np.concatenate([rng.random(5000) * 10.0 ** rng.integers(-300, 300, 5000),
                                 [0.0, np.inf, 1.0, np.nextafter(1.0, 2.0)]]))
    hw = [high_word(float(x)) for x in xs]
    assert all(a <= b for a, b in zip(hw, hw[1:]))
