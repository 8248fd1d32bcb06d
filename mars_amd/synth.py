"""Seeded synthetic greyscale and RGB test images for the search workloads.

The generator is this project's own (the reference ships none).  One plane is
the sum of
  * 3 linear gradients at uniformly random angles,
  * 4-octave smooth value noise (bilinear-smoothstep lattice, total amplitude 60),
  * 20 axis-aligned rectangles filled with a uniform random grey (overwrite),
  * Gaussian noise with sigma 4,
clipped to [0, 255] and rounded to u8.  Everything is drawn from one
``numpy.random.Generator(PCG64(seed))`` in a fixed order, so a seed names an
image; ``tests/test_synth.py`` pins the SHA-256 of the workload images.
"""

from __future__ import annotations

import hashlib

import numpy as np

__all__ = ["plane", "rgb", "sha256", "batch_specs", "WORKLOAD_SEEDS"]

# Seeds of the named workloads (256x256 fixed-grid, 512x512 quadtree,
# 768x512 RGB, the 64-image batch base, the 2048x2048 dense pool).
WORKLOAD_SEEDS = {"fixed256": 1, "quadtree512": 2, "rgb768x512": 3, "batch64": 100,
                  "dense2048": 7}


def _value_noise(rng: np.random.Generator, w: int, h: int, octaves: int = 4,
                 amplitude: float = 60.0) -> np.ndarray:
    out = np.zeros((h, w), dtype=np.float64)
    weights = [0.5 ** o for o in range(octaves)]
    norm = sum(weights)
    side = max(w, h)
    ys = np.arange(h, dtype=np.float64)
    xs = np.arange(w, dtype=np.float64)
    for o in range(octaves):
        cell = max(2.0, side / (4.0 * 2 ** o))
        gh = int(np.ceil(h / cell)) + 2
        gw = int(np.ceil(w / cell)) + 2
        lattice = rng.uniform(-1.0, 1.0, size=(gh, gw))
        fy = ys / cell
        fx = xs / cell
        iy = np.floor(fy).astype(np.int64)
        ix = np.floor(fx).astype(np.int64)
        ty = fy - iy
        tx = fx - ix
        sy = ty * ty * (3.0 - 2.0 * ty)
        sx = tx * tx * (3.0 - 2.0 * tx)
        a = lattice[iy][:, ix]
        b = lattice[iy][:, ix + 1]
        c = lattice[iy + 1][:, ix]
        d = lattice[iy + 1][:, ix + 1]
        top = a + (b - a) * sx[None, :]
        bot = c + (d - c) * sx[None, :]
        out += weights[o] / norm * (top + (bot - top) * sy[:, None])
    return amplitude * out


def _field(rng: np.random.Generator, w: int, h: int) -> np.ndarray:
    """Float image before clipping (mean near 128)."""
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    side = float(max(w, h))
    img = np.full((h, w), 128.0)
    for _ in range(3):
        theta = rng.uniform(0.0, 2.0 * np.pi)
        amp = rng.uniform(15.0, 45.0)
        proj = (xx - w / 2.0) * np.cos(theta) + (yy - h / 2.0) * np.sin(theta)
        img += amp * proj / side
    img += _value_noise(rng, w, h)
    for _ in range(20):
        rw = int(rng.integers(max(2, w // 32), max(3, w // 4) + 1))
        rh = int(rng.integers(max(2, h // 32), max(3, h // 4) + 1))
        x0 = int(rng.integers(0, max(1, w - rw + 1)))
        y0 = int(rng.integers(0, max(1, h - rh + 1)))
        grey = rng.uniform(0.0, 255.0)
        img[y0:y0 + rh, x0:x0 + rw] = grey
    img += rng.normal(0.0, 4.0, size=(h, w))
    return img


def _to_u8(img: np.ndarray) -> np.ndarray:
    return np.clip(np.rint(img), 0, 255).astype(np.uint8)


def plane(w: int, h: int, seed: int) -> np.ndarray:
    """One seeded ``h x w`` u8 greyscale plane (row-major, C-contiguous)."""
    rng = np.random.default_rng(seed)
    return np.ascontiguousarray(_to_u8(_field(rng, w, h)))


def rgb(w: int, h: int, seed: int) -> np.ndarray:
    """Seeded ``h x w x 3`` u8 RGB image whose channels correlate about 0.8.

    Channel c = 128 + (2 (G0 - 128) + (Gc - 128)) / sqrt(5) with G0 a shared
    field and Gc an independent one (correlation 4/5 for equal variances).
    """
    base = _field(np.random.default_rng(seed), w, h) - 128.0
    chans = []
    for c in range(3):
        own = _field(np.random.default_rng(seed * 1000 + 1 + c), w, h) - 128.0
        chans.append(_to_u8(128.0 + (2.0 * base + own) / np.sqrt(5.0)))
    return np.ascontiguousarray(np.stack(chans, axis=-1))


# The 64-image batch: sizes and RD lambdas drawn uniformly from these sets.
BATCH_SIZES = ((256, 256), (512, 512), (768, 512), (1024, 768))
BATCH_LAMBDAS = (50.0, 200.0, 800.0)


def batch_specs(n: int = 64, seed: int = WORKLOAD_SEEDS["batch64"]):
    """(w, h, plane seed, lambda) of each image of the seeded uneven batch.

    Sizes and lambdas come from one Generator(PCG64(seed)) stream; image i is
    ``plane(w, h, seed + i)``.
    """
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        w, h = BATCH_SIZES[int(rng.integers(len(BATCH_SIZES)))]
        lam = BATCH_LAMBDAS[int(rng.integers(len(BATCH_LAMBDAS)))]
        out.append((w, h, seed + i, lam))
    return out


def sha256(img: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(img).tobytes()).hexdigest()
