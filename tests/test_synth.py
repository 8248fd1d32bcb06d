"""Pins of the synthetic workload images (mars_amd/synth.py).

Every measured workload and every full-size parity test is named by a seed;
these SHA-256 values make the seed name the same pixels on every machine.  A
change of the generator or of numpy's PCG64 / distribution streams fails here
first instead of silently changing what the other tests and the bench measure.
"""

import hashlib
import json

import numpy as np

from mars_amd import synth

S = synth.WORKLOAD_SEEDS

# SHA-256 of the raw u8 pixels, row-major.
FIXED256 = "d0e45abd7c00e0497658d4240ca3420acced5499eb97d8e89e3f3d25f901fcad"
QUADTREE512 = "e05f66b454f1527de28d55e6d0406179dc6956b26e18d3fc3d8a27b3e8ea18e4"
RGB768X512 = "d8d2f3b1517436a8ece45909415f5296cb586eba1a0af70607ef940207f1158a"
RGB768X512_CHANNELS = (
    "7d868ca292e1c1348faea1a923fb56991791d4e824335f01023dbb32d9e4bf90",
    "1962f9a064fdf093f1c516f8b778143994a9b4cb703f014b66e3361ca4baf5d2",
    "57c6ba8f352c9007c788e4f3fc917c21e73f5eaa9fe9bc350e56d5325dab648b",
)
DENSE2048 = "47884df43ccdba2b5d44821be75017a51e858693949050143ff763c2bdc120c9"
# SHA-256 over the concatenated hex digests of the batch planes, in order.
BATCH64 = "d706ce3a965edf7ba4f7b91154e572ed8cd58735b0e22335eba4a7c9fe6e866d"
BATCH64_SPECS = "6a8324ece2fa608ee5332e09aba19bd027639eb292e1648ee48654ea7e496bec"
BATCH512 = "45c71b0557041cf0bebc6f2d0a8a8d2fb453c455dc7cf2cd4a213da80813857f"


def _digest_of_digests(planes):
    h = hashlib.sha256()
    for p in planes:
        h.update(synth.sha256(p).encode())
    return h.hexdigest()


def test_fixed256():
    img = synth.plane(256, 256, S["fixed256"])
    assert img.shape == (256, 256) and img.dtype == np.uint8 and img.flags.c_contiguous
    assert synth.sha256(img) == FIXED256


def test_quadtree512():
    assert synth.sha256(synth.plane(512, 512, S["quadtree512"])) == QUADTREE512


def test_rgb768x512_and_channels():
    rgb = synth.rgb(768, 512, S["rgb768x512"])
    assert rgb.shape == (512, 768, 3) and rgb.dtype == np.uint8
    assert synth.sha256(rgb) == RGB768X512
    assert tuple(synth.sha256(rgb[..., c]) for c in range(3)) == RGB768X512_CHANNELS
    # the stated inter-channel correlation (about 0.8)
    x = rgb.reshape(-1, 3).astype(np.float64).T
    cc = np.corrcoef(x)
    assert 0.6 < cc[0, 1] < 0.95 and 0.6 < cc[0, 2] < 0.95 and 0.6 < cc[1, 2] < 0.95


def test_dense2048():
    assert synth.sha256(synth.plane(2048, 2048, S["dense2048"])) == DENSE2048


def test_batch64_specs_and_planes():
    specs = synth.batch_specs()
    assert len(specs) == 64
    assert hashlib.sha256(json.dumps(specs).encode()).hexdigest() == BATCH64_SPECS
    assert {(w, h) for w, h, _, _ in specs} <= set(synth.BATCH_SIZES)
    assert {lam for *_, lam in specs} <= set(synth.BATCH_LAMBDAS)
    assert [seed for _, _, seed, _ in specs] == list(range(S["batch64"], S["batch64"] + 64))
    assert _digest_of_digests(synth.plane(w, h, seed) for w, h, seed, _ in specs) == BATCH64


def test_batch512():
    assert _digest_of_digests(synth.plane(512, 512, s) for s in range(200, 264)) == BATCH512


def test_seed_names_the_image():
    a = synth.plane(96, 64, 11)
    assert np.array_equal(a, synth.plane(96, 64, 11))
    assert not np.array_equal(a, synth.plane(96, 64, 12))
