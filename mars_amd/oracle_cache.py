"""The top-32 oracle cache and recall scoring, built on the GPU's ranked tables.

Mirrors the reference's bench crate:
  * crates/mars-bench/src/oracle.rs -- ``OracleConfig`` (51-86), ``manifest_hash``
    (128-140), ``CacheSlot`` (146-199, 28-byte records), ``SizeCache.block_at``
    (223-232), the ``MARSORC1`` layout ``serialize``/``parse`` (381-482), ``save``/``load``
    (483-505), ``load_and_validate`` (512-537) and ``build`` (539-575);
  * crates/mars-bench/src/recall.rs -- ``MethodPick`` (32-42), the JSONL
    ``read_picks``/``write_picks`` (63-97), ``oracle_top1_as_picks`` (100-126),
    ``regret_db`` (134-150), ``percentile_sorted`` (153-170) and ``score`` (234-280).

The reference fills the cache with its WGSL top-32 kernel, which ranks by a
binary32 fit.  ``build`` here ranks with the exact search (binary64 rms, ties in
enumeration order: ``search_top_k``, crates/mars-codec/src/encode.rs:676-788)
from a context prepared with ``set_ranked(32)``, and stores each rms rounded
to binary32 as the format's field.  The file layout is the reference's.
"""

from __future__ import annotations

import hashlib
import json
import math
import struct
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

CACHE_FORMAT_VERSION = 1
MAGIC = b"MARSORC1"
SLOTS = 32
RECORD_LEN = 28


class CacheError(Exception):
    """Base of the cache errors (oracle.rs:254-296)."""


class BadMagic(CacheError):
    pass


class VersionError(CacheError):
    pass


class Truncated(CacheError):
    pass


class Stale(CacheError):
    pass


@dataclass(frozen=True)
class OracleConfig:
    min_size: int
    max_size: int
    shift: int
    bits_alfa: int = 4
    bits_beta: int = 7
    max_alfa: float = 1.0

    def sizes(self) -> List[int]:
        """Every power-of-two block size from min_size up to max_size (oracle.rs:64-75)."""
        out, s = [], self.min_size
        while s <= self.max_size:
            out.append(s)
            s *= 2
            if s == 0:
                break
        return out


def manifest_hash(image_sha256_hex: str, config: OracleConfig) -> bytes:
    """oracle.rs:128-140: image hash, format version and every search parameter."""
    h = hashlib.sha256()
    h.update(image_sha256_hex.encode())
    h.update(struct.pack("<I", CACHE_FORMAT_VERSION))
    h.update(struct.pack("<5I", config.min_size, config.max_size, config.shift,
                         config.bits_alfa, config.bits_beta))
    h.update(struct.pack("<d", config.max_alfa))
    return h.digest()


@dataclass(frozen=True)
class CacheSlot:
    valid: bool = False
    dom_row: int = 0
    dom_col: int = 0
    isometry: int = 0
    qalfa: int = 0
    qbeta: int = 0
    rms: float = 0.0  # binary32 value

    def pack(self) -> bytes:
        return struct.pack("<6If", int(self.valid), self.dom_row, self.dom_col, self.isometry,
                           self.qalfa, self.qbeta, self.rms)

    @staticmethod
    def unpack(b: bytes) -> "CacheSlot":
        v, r, c, iso, qa, qb, rms = struct.unpack("<6If", b)
        return CacheSlot(v != 0, r, c, iso & 0xFF, qa, qb, rms)


INVALID = CacheSlot()


@dataclass
class SizeCache:
    size: int
    num_blocks_x: int
    num_blocks_y: int
    blocks: List[List[CacheSlot]]  # row-major, SLOTS each

    def block_at(self, row: int, col: int) -> Optional[List[CacheSlot]]:
        """oracle.rs:223-232: None when misaligned or outside the grid."""
        if row % self.size or col % self.size:
            return None
        bx, by = col // self.size, row // self.size
        if bx >= self.num_blocks_x or by >= self.num_blocks_y:
            return None
        return self.blocks[by * self.num_blocks_x + bx]


@dataclass
class OracleCache:
    image_name: str
    image_sha256_hex: str
    config: OracleConfig
    sizes: List[SizeCache] = field(default_factory=list)

    def size(self, s: int) -> Optional[SizeCache]:
        return next((x for x in self.sizes if x.size == s), None)


def _string(s: str) -> bytes:
    b = s.encode()
    return struct.pack("<I", len(b)) + b


def serialize(cache: OracleCache) -> bytes:
    """The MARSORC1 layout of oracle.rs:368-407, little-endian throughout."""
    c = cache.config
    out = bytearray(MAGIC)
    out += struct.pack("<I", CACHE_FORMAT_VERSION)
    out += manifest_hash(cache.image_sha256_hex, c)
    out += _string(cache.image_name) + _string(cache.image_sha256_hex)
    out += struct.pack("<5I", c.min_size, c.max_size, c.shift, c.bits_alfa, c.bits_beta)
    out += struct.pack("<d", c.max_alfa)
    out += struct.pack("<I", len(cache.sizes))
    for sc in cache.sizes:
        out += struct.pack("<3I", sc.size, sc.num_blocks_x, sc.num_blocks_y)
        for block in sc.blocks:
            for slot in block:
                out += slot.pack()
    return bytes(out)


class _Cursor:
    def __init__(self, b: bytes):
        self.b, self.pos = b, 0

    def take(self, n: int) -> bytes:
        if self.pos + n > len(self.b):
            raise Truncated(f"expected {n} more bytes at offset {self.pos}, "
                            f"only {len(self.b) - self.pos} remain")
        s = self.b[self.pos:self.pos + n]
        self.pos += n
        return s

    def u32(self) -> int:
        return struct.unpack("<I", self.take(4))[0]

    def string(self) -> str:
        raw = self.take(self.u32())
        try:
            return raw.decode()
        except UnicodeDecodeError as e:
            raise Truncated(f"non-utf8 string: {e}") from None


def parse(b: bytes) -> OracleCache:
    """oracle.rs:412-482: magic, version, then a header whose stored hash must
    match its own fields (a mismatch is corruption, not staleness)."""
    cur = _Cursor(b)
    if cur.take(8) != MAGIC:
        raise BadMagic("not a mars oracle cache file (bad magic bytes)")
    version = cur.u32()
    if version != CACHE_FORMAT_VERSION:
        raise VersionError(f"unsupported cache format version {version}")
    stored = cur.take(32)
    name, sha = cur.string(), cur.string()
    mn, mx, sh, ba, bb = (cur.u32() for _ in range(5))
    ma = struct.unpack("<d", cur.take(8))[0]
    config = OracleConfig(mn, mx, sh, ba, bb, ma)
    if manifest_hash(sha, config) != stored:
        raise Truncated("manifest hash does not match the header's own stored fields")
    sizes = []
    for _ in range(cur.u32()):
        size, nx, ny = cur.u32(), cur.u32(), cur.u32()
        blocks = [[CacheSlot.unpack(cur.take(RECORD_LEN)) for _ in range(SLOTS)]
                  for _ in range(nx * ny)]
        sizes.append(SizeCache(size, nx, ny, blocks))
    return OracleCache(name, sha, config, sizes)


def save(cache: OracleCache, path) -> None:
    with open(path, "wb") as f:
        f.write(serialize(cache))


def load(path) -> OracleCache:
    """A cache file parsed with no external validation (oracle.rs:497-505)."""
    with open(path, "rb") as f:
        return parse(f.read())


def load_and_validate(path, image_sha256_hex: str, config: OracleConfig) -> OracleCache:
    """oracle.rs:512-537: a cache built from another image or config is an error,
    never silently rebuilt or ignored."""
    cache = load(path)
    if manifest_hash(image_sha256_hex, config) != manifest_hash(cache.image_sha256_hex,
                                                                cache.config):
        raise Stale(f"STALE ORACLE CACHE {path}: stored image={cache.image_name} "
                    f"sha256={cache.image_sha256_hex} config={cache.config} | current "
                    f"sha256={image_sha256_hex} config={config}")
    return cache


def _slot(c) -> CacheSlot:
    return CacheSlot(True, c.dom_row, c.dom_col, c.isometry, c.qalfa, c.qbeta,
                     float(np.float32(c.rms)))


def assemble(image: np.ndarray, image_name: str, config: OracleConfig, lookup,
             image_sha256_hex: Optional[str] = None) -> OracleCache:
    """The cache from `lookup(row, col, size) -> ascending candidate list`,
    every size's full grid of aligned blocks, row-major (oracle.rs:539-575).
    The image hash defaults to SHA-256 of the plane's row-major bytes; the
    reference hashes the image file, so pass that hex digest to match it."""
    h, w = image.shape
    if image_sha256_hex is None:
        image_sha256_hex = hashlib.sha256(np.ascontiguousarray(image).tobytes()).hexdigest()
    cache = OracleCache(image_name, image_sha256_hex, config)
    for size in config.sizes():
        nx, ny = w // size, h // size
        blocks = []
        for by in range(ny):
            for bx in range(nx):
                lst = list(lookup(by * size, bx * size, size))[:SLOTS]
                blocks.append([_slot(c) for c in lst] + [INVALID] * (SLOTS - len(lst)))
        cache.sizes.append(SizeCache(size, nx, ny, blocks))
    return cache


def build(ctx, image: np.ndarray, image_name: str, config: OracleConfig,
          image_sha256_hex: Optional[str] = None) -> OracleCache:
    """The cache of one image from one ranked (k = 32) step on the GPU.

    Every aligned grid block of a size in [min_size, max_size] lying wholly
    inside the plane is a node of the quadtree walk, so the prepared tables
    hold all of them."""
    from .provider import EncodeParams

    params = EncodeParams(min_size=config.min_size, max_size=config.max_size,
                          shift=config.shift, bits_alfa=config.bits_alfa,
                          bits_beta=config.bits_beta, max_alfa=config.max_alfa)
    ctx.set_ranked(SLOTS)
    ctx.prepare(image, params, [config.shift])
    index = {}
    for (size, shift), (rows, cols, lists) in ctx.ranked_tables().items():
        if shift == config.shift:
            index.update({(size, r, c): l for r, c, l in zip(rows, cols, lists)})

    def lookup(row, col, size):
        try:
            return index[(size, row, col)]
        except KeyError:
            raise RuntimeError(f"no ranked table entry for block ({row},{col}) size {size}") \
                from None

    return assemble(image, image_name, config, lookup, image_sha256_hex)


# ---- recall (recall.rs) -------------------------------------------------------

@dataclass(frozen=True)
class MethodPick:
    row: int
    col: int
    size: int
    dom_row: int
    dom_col: int
    isometry: int
    qalfa: int
    qbeta: int
    rms: float

    def to_json(self) -> str:
        return json.dumps(self.__dict__)

    @staticmethod
    def from_json(line: str) -> "MethodPick":
        return MethodPick(**json.loads(line))


def write_picks(path, picks: Sequence[MethodPick]) -> None:
    """One JSON object per line (recall.rs:63-97)."""
    with open(path, "w") as f:
        for pk in picks:
            f.write(pk.to_json() + "\n")


def read_picks(path) -> List[MethodPick]:
    with open(path) as f:
        return [MethodPick.from_json(line) for line in f if line.strip()]


def oracle_top1_as_picks(cache: OracleCache, size: int) -> List[MethodPick]:
    """recall.rs:100-126: every block's rank-0 entry as a pick."""
    sc = cache.size(size)
    if sc is None:
        return []
    out = []
    for by in range(sc.num_blocks_y):
        for bx in range(sc.num_blocks_x):
            best = sc.blocks[by * sc.num_blocks_x + bx][0]
            if best.valid:
                out.append(MethodPick(by * size, bx * size, size, best.dom_row, best.dom_col,
                                      best.isometry, best.qalfa, best.qbeta, float(best.rms)))
    return out


def regret_db(rms_method: float, rms_oracle_best: float) -> Optional[float]:
    """recall.rs:134-150: 20 log10(method / oracle), None unless both are > 0."""
    if (math.isnan(rms_method) or math.isnan(rms_oracle_best) or rms_oracle_best <= 0.0
            or rms_method <= 0.0):
        return None
    return 20.0 * math.log10(rms_method / rms_oracle_best)


def percentile_sorted(sorted_vals: Sequence[float], p: float) -> float:
    """recall.rs:153-170: linear interpolation between closest ranks."""
    if not sorted_vals:
        raise ValueError("percentile of an empty sample")
    if len(sorted_vals) == 1:
        return sorted_vals[0]
    rank = p / 100.0 * (len(sorted_vals) - 1)
    lo, hi = math.floor(rank), math.ceil(rank)
    if lo == hi:
        return sorted_vals[lo]
    frac = rank - lo
    return sorted_vals[lo] * (1.0 - frac) + sorted_vals[hi] * frac


@dataclass
class RecallReport:
    matched: int = 0
    unmatched: int = 0
    top1: int = 0
    top5: int = 0
    top32: int = 0
    regret_db_values: List[float] = field(default_factory=list)

    def rate(self, hits: int) -> float:
        return hits / self.matched if self.matched else 0.0

    def top1_rate(self) -> float:
        return self.rate(self.top1)

    def top5_rate(self) -> float:
        return self.rate(self.top5)

    def top32_rate(self) -> float:
        return self.rate(self.top32)

    def regret_stats(self):
        if not self.regret_db_values:
            return None
        v = sorted(self.regret_db_values)
        return {"mean_db": sum(v) / len(v), "median_db": percentile_sorted(v, 50.0),
                "p95_db": percentile_sorted(v, 95.0), "n": len(v)}


def score(cache: OracleCache, picks: Sequence[MethodPick]) -> RecallReport:
    """recall.rs:234-280: exact-tuple rank of each pick in its block's list,
    and its rms regret against the block's rank-0 entry."""
    rep = RecallReport()
    for pk in picks:
        sc = cache.size(pk.size)
        block = sc.block_at(pk.row, pk.col) if sc else None
        if block is None or not block[0].valid:
            rep.unmatched += 1
            continue
        rep.matched += 1
        rank = next((i for i, c in enumerate(block)
                     if c.valid and (c.dom_row, c.dom_col, c.isometry, c.qalfa, c.qbeta)
                     == (pk.dom_row, pk.dom_col, pk.isometry, pk.qalfa, pk.qbeta)), None)
        if rank is not None:
            rep.top1 += rank == 0
            rep.top5 += rank < 5
            rep.top32 += rank < 32
        r = regret_db(pk.rms, float(block[0].rms))
        if r is not None:
            rep.regret_db_values.append(r)
    return rep
