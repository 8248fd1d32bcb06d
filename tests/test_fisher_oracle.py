"""CPU pins of the oracle's Fisher search and colour plumbing.

What fixes each part independently of the restatement in oracle/fisher.c:
  * ORDERING / MAPPING tables, newclass, variance_class, flips and the bucket
    index -- the reference's own 1998 C codec (reference/mars1, built
    unchanged into oracle/_ref by `make -C oracle ref`): tables read from its
    globals, classification compared on random blocks (tie-heavy ones
    included), FisherIndexing's buckets compared as sets (its lists are in
    reverse scan order; the Rust port pushes in scan order);
  * MAPPING also by its meaning: a range that is an isometric image of a
    domain is mapped back onto it exactly by permute_range with isometry
    MAPPING[range iso][domain iso];
  * the search over a bucket -- brute force in numpy on small planes, the
    exhaustive search as a lower bound, and equality with it whenever the
    exhaustive winner is among the bucket's candidates;
  * the device classification (mars_amd/csrc/fisher_class.h) -- compiled
    for the host and compared with the oracle block by block;
  * ycbcr / downsample_box / the MARC container -- the reference's own
    test cases (metrics.rs:553-600, color.rs:866-881), exact rational
    arithmetic, and the byte layout of color.rs:778-790.
"""

import ctypes as C
import subprocess
from fractions import Fraction

import numpy as np
import pytest

from mars_amd import codec, synth
from tests import _mars1 as M
from tests import _oracle as O

needs_ref = pytest.mark.skipif(not M.available(), reason="reference C sources not present")

SIZES = (2, 4, 8, 16, 32)


def _blocks(rng, size, n):
    """Random D-scaled blocks: many ties (values 0..1, 0..3), general ones and flat ones."""
    for t in range(n):
        hi = (2, 4, 1021, 1)[t % 4]
        yield rng.integers(0, hi, (size, size)).astype(np.float64) + (500.0 if hi == 1 else 0.0)


@needs_ref
def test_tables_equal_the_reference_c_tables():
    o, m = O.fisher_tables()
    ro, rm = M.tables()
    assert np.array_equal(o, ro)
    assert np.array_equal(m, rm)


@needs_ref
@pytest.mark.parametrize("size", SIZES)
def test_classification_equals_the_reference_c(size):
    rng = np.random.default_rng(size)
    for b in _blocks(rng, size, 1200):
        assert O.fisher_newclass(b) == M.newclass(b)
        assert O.fisher_variance_class(b) == M.variance_class(b)
        # The reference classifies 2x2 means, the Rust port (and the oracle) sums.
        assert O.fisher_newclass(b) == M.newclass(b / 4.0)
        assert O.fisher_variance_class(b) == M.variance_class(b / 4.0)
    b = rng.integers(0, 1021, (size, size)).astype(np.float64)
    for iso in range(8):
        assert np.array_equal(O.fisher_flips(b, iso), M.flips(b, iso))


@needs_ref
@pytest.mark.parametrize("w,h,size,shift,seed", [
    (96, 64, 4, 4, 11), (96, 64, 8, 4, 11), (96, 64, 2, 2, 11), (75, 53, 16, 8, 15),
    (64, 64, 4, 1, 12), (20, 12, 4, 2, 14), (40, 40, 16, 4, 3), (12, 12, 4, 4, 1),
])
def test_buckets_equal_the_reference_c_indexing(w, h, size, shift, seed):
    img = synth.plane(w, h, seed)
    ref = M.fisher_index(O.contract(img).astype(np.float64), w, h, size, shift)
    mine, own = O.fisher_index(img, size, shift)
    assert [sorted(b) for b in mine] == [sorted(b) for b in ref]
    P = ((h - 2 * size) // shift + 1) * ((w - 2 * size) // shift + 1) if min(w, h) >= 2 * size else 0
    assert sum(own) == P
    nonempty = [b for b in range(72) if own[b]]
    for b in range(72):  # empty buckets alias the first non-empty one
        if not own[b] and nonempty:
            assert mine[b] == mine[nonempty[0]]


def test_buckets_of_a_plane_without_positions_are_empty():
    img = synth.plane(12, 12, 1)
    mine, own = O.fisher_index(img, 8, 4)
    assert sum(own) == 0 and all(not b for b in mine)
    res, each = O.fisher_table(img, 8, 4, [0], [0])
    assert res == [None] and each == [0]


def _inv_flip(x, iso):
    s = x.shape[0]
    src = O.fisher_flips(np.arange(s * s, dtype=np.float64).reshape(s, s), iso).astype(int)
    d = np.zeros(s * s)
    d[src.ravel()] = x.ravel()
    return d.reshape(s, s)


@pytest.mark.parametrize("size", (2, 4, 8))
def test_mapping_maps_an_isometric_range_back_onto_its_domain(size):
    """flips(D, b) == flips(R, a)  =>  permute_range(R, MAPPING[a][b]) == D."""
    _, m = O.fisher_tables()
    x = np.random.default_rng(5).permutation(size * size).reshape(size, size).astype(np.float64)
    for a in range(8):
        for b in range(8):
            d, r = _inv_flip(x, b), _inv_flip(x, a)
            got = O.permute_range(r.astype(np.uint8), int(m[a][b])).astype(np.float64)
            assert np.array_equal(got, d), (a, b)


def _brute(img, size, shift, row, col):
    """search_block_fitted over the Fisher bucket, written out in numpy."""
    h, w = img.shape
    d = O.contract(img).astype(np.int64)
    _, mapping = O.fisher_tables()
    P = []
    if h >= 2 * size and w >= 2 * size:
        P = [(r, c) for r in range(0, h - 2 * size + 1, shift)
             for c in range(0, w - 2 * size + 1, shift)]
    buckets = {}
    for r, c in P:
        blk = d[r // 2:r // 2 + size, c // 2:c // 2 + size].astype(np.float64)
        iso, clas, vc = O.fisher_classify(blk)
        buckets.setdefault(clas * 24 + vc, []).append((r, c, iso))
    first = min(buckets) if buckets else None
    rng_blk = img[row:row + size, col:col + size]
    isom, clas, vc = O.fisher_classify(rng_blk.astype(np.float64))
    cands = buckets.get(clas * 24 + vc, buckets.get(first, []))
    t0 = int(rng_blk.astype(np.int64).sum())
    t2 = int((rng_blk.astype(np.int64) ** 2).sum())
    best = None
    for r, c, diso in cands:
        k = int(mapping[isom][diso])
        win = d[r // 2:r // 2 + size, c // 2:c // 2 + size]
        t1 = int((win * O.permute_range(rng_blk, k).astype(np.int64)).sum())
        m = (size * size, int(win.sum()), int((win * win).sum()), t0, t1, t2)
        qa, qb, rms = O.fit_f64(m)
        if best is None or rms < best[0]:
            best = (rms, r, c, k, qa, qb, m)
    return best, len(cands)


@pytest.mark.parametrize("w,h,size,shift,seed", [
    (40, 36, 4, 4, 21), (40, 36, 8, 4, 22), (33, 31, 2, 2, 23), (48, 48, 16, 8, 24),
])
def test_search_equals_brute_force(w, h, size, shift, seed):
    img = synth.plane(w, h, seed)
    rows = [r for r in range(0, h - size + 1, size)]
    cols = [c for c in range(0, w - size + 1, size)]
    rr = [r for r in rows for _ in cols]
    cc = [c for _ in rows for c in cols]
    res, each = O.fisher_table(img, size, shift, rr, cc, threads=4)
    for r, c, got, ev in zip(rr, cc, res, each):
        best, n = _brute(img, size, shift, r, c)
        assert ev == n
        if best is None:
            assert got is None
            continue
        rms, dr, dc, k, qa, qb, m = best
        assert (got.dom_row, got.dom_col, got.isometry, got.qalfa, got.qbeta) == (dr, dc, k, qa, qb)
        assert got.rms == rms
        assert (got.m.s0, got.m.s1_x4, got.m.s2_x16, got.m.t0, got.m.t1_x4, got.m.t2) == m


def test_fisher_is_bounded_by_the_exhaustive_search():
    img = synth.plane(96, 64, 31)
    for size in (4, 8, 16):
        nodes = O.nodes(96, 64, size, size)[size]
        rows, cols = [r for r, _ in nodes], [c for _, c in nodes]
        fis, each = O.fisher_table(img, size, 4, rows, cols, threads=8)
        exh, _ = O.search_table(img, size, 4, rows, cols, threads=8)
        buckets, _ = O.fisher_index(img, size, 4)
        hits = 0
        for f, e, n in zip(fis, exh, each):
            assert f is not None and e is not None and n > 0
            assert f.rms >= e.rms
            cand = (e.dom_row, e.dom_col, e.isometry)
            if f.rms == e.rms:
                hits += 1
            if cand == (f.dom_row, f.dom_col, f.isometry):
                assert f.rms == e.rms
        assert hits > 0


def _header_classifier(tmp_path):
    """fisher_class.h compiled for the host (the code the GPU kernels run)."""
    src = tmp_path / "fc.cpp"
    src.write_text(
        '#define MARS_HD\n#include "fisher_class.h"\n'
        'extern "C" unsigned fc(const long long *s, const long long *q, unsigned h) {\n'
        '  int64_t a[4] = {s[0], s[1], s[2], s[3]}, b[4] = {q[0], q[1], q[2], q[3]};\n'
        '  return mars::fisher::classify(a, b, h); }\n')
    so = tmp_path / "fc.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-shared", "-fPIC",
                    f"-I{O.ROOT / 'mars_amd' / 'csrc'}", str(src), "-o", str(so)], check=True)
    L = C.CDLL(str(so))
    L.fc.argtypes = [C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), C.c_uint]
    L.fc.restype = C.c_uint
    return L


@pytest.mark.parametrize("size", SIZES)
def test_device_classifier_equals_the_oracle(tmp_path, size):
    L = _header_classifier(tmp_path)
    h = size // 2
    rng = np.random.default_rng(100 + size)
    for b in _blocks(rng, size, 800):
        q = [b[i * h:(i + 1) * h, j * h:(j + 1) * h].astype(np.int64) for i in (0, 1) for j in (0, 1)]
        s = (C.c_longlong * 4)(*[int(x.sum()) for x in q])
        sq = (C.c_longlong * 4)(*[int((x * x).sum()) for x in q])
        iso, clas, vc = O.fisher_classify(b)
        assert L.fc(s, sq, h) == (clas * 24 + vc) * 8 + iso


# ---- colour --------------------------------------------------------------

def _round_half_away(x: Fraction) -> int:
    n = abs(x)
    r = int(n + Fraction(1, 2)) if n.denominator != 1 else int(n)
    return r if x >= 0 else -r


def _exact_ycbcr(r, g, b):
    F = Fraction
    y = F("0.299") * r + F("0.587") * g + F("0.114") * b
    cb = 128 - F("0.168736") * r - F("0.331264") * g + F("0.5") * b
    cr = 128 + F("0.5") * r - F("0.418688") * g - F("0.081312") * b
    return [min(255, max(0, _round_half_away(v))) for v in (y, cb, cr)], (y, cb, cr)


def test_ycbcr_known_answers_and_exact_arithmetic():
    # metrics.rs:553-569: grey input gives Cb == Cr == 128 and Y == the grey value.
    grey = np.repeat(np.arange(256, dtype=np.uint8)[None, :, None], 3, axis=2)
    y, cb, cr = O.ycbcr(grey)
    assert np.array_equal(y[0], np.arange(256)) and (cb == 128).all() and (cr == 128).all()
    # JPEG's full-range primaries.
    prim = np.array([[[255, 0, 0], [0, 255, 0], [0, 0, 255], [255, 255, 255], [0, 0, 0]]],
                    np.uint8)
    y, cb, cr = O.ycbcr(prim)
    assert list(zip(y[0], cb[0], cr[0])) == [(76, 85, 255), (150, 44, 21), (29, 255, 107),
                                             (255, 128, 128), (0, 128, 128)]
    rng = np.random.default_rng(7)
    px = rng.integers(0, 256, (1, 4000, 3), dtype=np.uint8)
    y, cb, cr = O.ycbcr(px)
    for i in range(px.shape[1]):
        want, exact = _exact_ycbcr(*[int(v) for v in px[0, i]])
        for got, w_, e in zip((y[0, i], cb[0, i], cr[0, i]), want, exact):
            # binary64 may round the other way only right at a .5 boundary
            frac = abs(e - int(e)) if e >= 0 else abs(-e - int(-e))
            if abs(frac - Fraction(1, 2)) > Fraction(1, 10 ** 9):
                assert got == w_, (px[0, i], got, w_)


def test_ycbcr_round_trip_is_within_two_levels():
    """metrics.rs:571-600: RGB -> YCbCr -> RGB within 2 LSB (inverse of metrics.rs:434-460)."""
    n = 64
    x, yy = np.meshgrid(np.arange(n), np.arange(n))
    rgb = np.stack([(x * 4) % 256, (yy * 4) % 256, ((x + yy) * 3) % 256], axis=-1).astype(np.uint8)
    y, cb, cr = O.ycbcr(rgb)
    Y, Cb, Cr = y.astype(np.float64), cb.astype(np.float64) - 128.0, cr.astype(np.float64) - 128.0

    def u8(v):
        return np.clip(np.where(v >= 0, np.floor(v + 0.5), -np.floor(-v + 0.5)), 0, 255)

    back = np.stack([u8(Y + 1.402 * Cr), u8(Y - 0.344136 * Cb - 0.714136 * Cr),
                     u8(Y + 1.772 * Cb)], axis=-1)
    assert np.abs(back - rgb.astype(np.float64)).max() <= 2


def test_product_ycbcr_equals_the_oracle():
    rgb = synth.rgb(67, 45, 3)
    assert all(np.array_equal(a, b) for a, b in zip(codec.ycbcr(rgb), O.ycbcr(rgb)))


@pytest.mark.parametrize("w,h", [(5, 3), (8, 8), (1, 1), (7, 1), (1, 6), (384, 255)])
def test_downsample_box(w, h):
    p = synth.plane(max(w, 2), max(h, 2), w * 31 + h)[:h, :w]
    got = O.downsample_box(p)
    assert got.shape == ((h + 1) // 2, (w + 1) // 2)  # color.rs:866-874
    e = np.pad(p.astype(np.int64), ((0, h % 2), (0, w % 2)), mode="edge")
    want = (e[0::2, 0::2] + e[0::2, 1::2] + e[1::2, 0::2] + e[1::2, 1::2] + 2) // 4
    assert np.array_equal(got, want)
    assert np.array_equal(codec.downsample_box(p), got)
    flat = np.full((h, w), 100, np.uint8)  # color.rs:876-881
    assert (O.downsample_box(flat) == 100).all()


def test_container_layout():
    streams = [b"abc", b"", b"\x01" * 300]
    out = codec.write_container(codec.COLOR_RGB_420, streams)
    want = b"MARC" + bytes([0, 2, 3])
    for s in streams:
        want += len(s).to_bytes(4, "little") + s
    assert out == want
    assert codec.wrap_gray(b"xyz") == codec.write_container(codec.COLOR_GRAY, [b"xyz"])
