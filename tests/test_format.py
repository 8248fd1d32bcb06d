"""The stream formats read back on the CPU: `.mars` v0 (mars_format.rs:302-388),
the legacy Mars 1 `.ifs` bitstream (ifs.rs:183-278) and the iterative decoder
(ifs.rs:281-309, 406-484).

Pins:
  * 36 of the reference's golden `.ifs` fixtures (tests/golden/mars1, copied
    with their manifest entries by tests/golden/make_mars1_subset.py): every
    one parses to the transform and zero-alfa counts the 1998 encoder printed
    and decodes (10 iterations) to the SHA-256 of `decmars -i`'s own PGM;
    the two worked examples the reference's tests spell out (flat128: every
    pixel 129; tiny64: the section 12 leaf table).
  * `.mars` round trips of real encodes (threshold walk, RD with all four
    leaf modes incl. mode 3 residuals, adaptive density, lambda-adaptive step,
    odd sizes): read(bytes) re-written is the same bytes, the leaves tile the
    plane, the header carries the encode's parameters.
  * The reference's own reader tests (mars_format.rs:867-906) and every other
    typed error the readers return.
"""

import hashlib
import json
import struct
from pathlib import Path

import numpy as np
import pytest

from mars_amd import codec, synth
from mars_amd.codec import EncodeOptions, FormatError, encode_plane
from mars_amd.provider import EncodeParams
from tests import _oracle as O

GOLDEN = Path(__file__).resolve().parent / "golden" / "mars1"
MANIFEST = json.loads((GOLDEN / "manifest.json").read_text())


def _pgm_sha(img: np.ndarray) -> str:
    h, w = img.shape
    return hashlib.sha256(f"P5\n{w} {h}\n255\n".encode() + img.tobytes()).hexdigest()


# ---------------------------------------------------------------- .ifs + decoder

@pytest.mark.parametrize("stem", sorted(MANIFEST))
def test_golden_ifs_fixture_parses_and_decodes(stem):
    rec = MANIFEST[stem]
    data = (GOLDEN / rec["file"]).read_bytes()
    assert hashlib.sha256(data).hexdigest() == rec["ifs_sha256"]
    hdr, leaves = codec.ifs_parse(data)
    assert (hdr["width"], hdr["height"]) == (rec["width"], rec["height"])
    assert len(leaves) == rec["transforms"]
    assert sum(1 for l in leaves if l[4] == 0) == rec["zero_alfa_transforms"]
    img = codec.decode_iterative(hdr, leaves, 10)
    assert _pgm_sha(img) == rec["decode_iterative_sha256"]


def test_flat128_worked_example():
    """crates/mars-codec/tests/ifs.rs:26-52: 256 DC leaves of size 16, qbeta 64,
    every pixel decodes to 129."""
    hdr, leaves = codec.ifs_parse((GOLDEN / "flat128__default__fisher__r8.ifs").read_bytes())
    assert len(leaves) == 256
    assert all(l[4] == 0 and l[2] == 16 and l[5] == 64 for l in leaves)
    assert (codec.decode_iterative(hdr, leaves, 10) == 129).all()


def test_tiny64_walkthrough_table():
    """crates/mars-codec/tests/ifs.rs:55-110: the section 12 leaf table."""
    hdr, leaves = codec.ifs_parse((GOLDEN / "tiny64__walkthrough__fisher__r8.ifs").read_bytes())
    assert (hdr["min_size"], hdr["max_size"], hdr["width"], hdr["height"]) == (16, 32, 64, 64)
    expected = [(0, 0, 32, 0, 50, 0, 0, 0), (32, 0, 16, 8, 49, 0, 32, 0),
                (48, 0, 16, 8, 70, 0, 32, 0), (32, 16, 16, 8, 49, 0, 32, 0),
                (48, 16, 16, 8, 70, 0, 32, 0), (0, 32, 32, 0, 85, 0, 0, 0),
                (32, 32, 32, 0, 30, 0, 0, 0)]
    assert len(leaves) == 7
    for l, (row, col, size, qa, qb, iso, dr, dc) in zip(leaves, expected):
        assert (l[0], l[1], l[2], l[4], l[5]) == (row, col, size, qa, qb)
        if qa:
            assert (l[6], l[7], l[8]) == (iso, dr, dc)
    assert _pgm_sha(codec.decode_iterative(hdr, leaves, 10)) == \
        "341a36ae1b0938650ecee68024c750773534ce9f1db50bf6bc3b4c777ba23f44"


def test_ifs_rejects_damaged_streams():
    data = (GOLDEN / "tiny64__walkthrough__fisher__r8.ifs").read_bytes()
    with pytest.raises(FormatError) as e:
        codec.ifs_parse(data[:5])
    assert e.value.code == codec.FMT_E_TRUNCATED
    with pytest.raises(FormatError) as e:
        codec.ifs_parse(data + b"\0")
    assert e.value.code == codec.FMT_E_TRAILING_BITS
    # this stream ends with padding bits (ifs.rs:205-214): they must be zero
    with pytest.raises(FormatError) as e:
        codec.ifs_parse(data[:-1] + bytes([data[-1] | 1]))
    assert e.value.code == codec.FMT_E_NONZERO_PADDING
    # width 0: the 12-bit field after 4+4+7+7+6 = 28 header bits
    bad = bytearray(data)
    for bit in range(28, 40):
        bad[bit >> 3] &= ~(0x80 >> (bit & 7)) & 0xFF
    with pytest.raises(FormatError) as e:
        codec.ifs_parse(bytes(bad))
    assert e.value.code == codec.FMT_E_DEGENERATE_HEADER


def test_decode_refuses_leaves_outside_the_plane():
    hdr, leaves = codec.ifs_parse((GOLDEN / "tiny64__walkthrough__fisher__r8.ifs").read_bytes())
    bad = list(leaves)
    l = list(bad[1])
    l[7] = 64  # dom_row past the plane
    bad[1] = tuple(l)
    with pytest.raises(ValueError):
        codec.decode_iterative(hdr, bad, 1)


# ---------------------------------------------------------------- .mars round trips

def _textured_gradient(w, h):
    """mars_format.rs:805-818: a vertical gradient plus local texture (every mode)."""
    r, c = np.mgrid[0:h, 0:w]
    base = r * 255 // max(h, 1)
    texture = ((r // 4) * 13 + (c // 4) * 7) % 32 - 16
    return np.clip(base + texture, 0, 255).astype(np.uint8)


def _encode(img, p, o):
    prov = O.Provider(img, p.max_alfa, p.bits_alfa, p.bits_beta)
    data, _ = encode_plane(img, p, o, prov.fn, prov.user)
    return data


ROUND_TRIPS = [
    ("threshold walk", lambda: synth.plane(64, 48, 1),
     EncodeParams(min_size=4, max_size=16, shift=4), EncodeOptions()),
    ("RD all modes", lambda: _textured_gradient(96, 96),
     EncodeParams(min_size=4, max_size=16, shift=4, lambda_=40.0), EncodeOptions()),
    ("RD adaptive density + lambda step", lambda: synth.plane(80, 64, 2),
     EncodeParams(min_size=4, max_size=16, shift=4, lambda_=200.0),
     EncodeOptions(adaptive_density=True, lambda_adaptive_qstep=True)),
    ("odd sides, size 2 borders", lambda: synth.plane(75, 53, 91),
     EncodeParams(min_size=2, max_size=16, shift=2, lambda_=120.0, max_alfa=0.75),
     EncodeOptions(fixed_qstep=5.5)),
    ("quadtree512 lambda 200", lambda: synth.plane(128, 128, 2),
     EncodeParams(min_size=4, max_size=32, shift=4, lambda_=200.0),
     EncodeOptions(allowed_modes=(True, False, True, False))),
]


@pytest.mark.parametrize("case", range(len(ROUND_TRIPS)), ids=[c[0] for c in ROUND_TRIPS])
def test_mars_round_trip_of_a_real_encode(case):
    name, make, p, o = ROUND_TRIPS[case]
    img = make()
    h, w = img.shape
    data = _encode(img, p, o)
    hdr, leaves = codec.read_mars(data)
    assert (hdr["width"], hdr["height"], hdr["min_size"], hdr["max_size"], hdr["shift"],
            hdr["bits_alfa"], hdr["bits_beta"]) == (w, h, p.min_size, p.max_size, p.shift,
                                                    p.bits_alfa, p.bits_beta)
    assert codec.write_mars(hdr, leaves) == data
    # the leaves tile the plane exactly once
    cover = np.zeros((h, w), np.int32)
    for l in leaves:
        cover[l[0]:l[0] + l[2], l[1]:l[1] + l[2]] += 1
        assert l[0] + l[2] <= h and l[1] + l[2] <= w
        if l[3] >= 2:
            assert l[4] >= 1 and l[7] + 2 * l[2] <= h and l[8] + 2 * l[2] <= w
            assert len(l[11]) == (l[2] * l[2] if l[3] == 3 else 0)
    assert (cover == 1).all()
    # decoding reconstructs the plane closely (a property, not a pin)
    dec = codec.decode_iterative(hdr, leaves, 10).astype(np.float64)
    mse = float(((dec - img.astype(np.float64)) ** 2).mean())
    psnr = 10 * np.log10(255.0 ** 2 / max(mse, 1e-12))
    assert psnr > 20.0, (name, psnr)
    if name == "RD all modes":
        assert len({l[3] for l in leaves}) > 1


def test_every_mode_round_trips():
    """mars_format.rs:793-842: all modes the RD search picks, including mode 3's
    residual levels, come back field for field."""
    img = _textured_gradient(96, 96)
    data = _encode(img, EncodeParams(min_size=4, max_size=16, shift=4, lambda_=40.0),
                   EncodeOptions())
    hdr, leaves = codec.read_mars(data)
    modes = {l[3] for l in leaves}
    assert len(modes) > 1
    hdr2, leaves2 = codec.read_mars(codec.write_mars(hdr, leaves))
    assert hdr2 == hdr and leaves2 == leaves


# ---------------------------------------------------------------- typed errors

def _small_stream():
    return _encode(synth.plane(32, 32, 3), EncodeParams(min_size=4, max_size=16, shift=4),
                   EncodeOptions())


def _code(data):
    with pytest.raises(FormatError) as e:
        codec.read_mars(bytes(data))
    return e.value.code


def test_rejects_bad_magic():
    b = bytearray(29)
    b[0:4] = b"NOPE"
    assert _code(b) == codec.FMT_E_BAD_MAGIC


def test_rejects_truncated_input():
    assert _code(b"") == codec.FMT_E_TRUNCATED
    assert _code(b"MARS") == codec.FMT_E_TRUNCATED
    data = _small_stream()
    assert _code(data[:26]) == codec.FMT_E_TRUNCATED  # section table entry cut


def test_rejects_oversized_section_claims():
    b = bytearray(_small_stream())
    b[25:29] = struct.pack("<I", 0xFFFFFFFF)
    assert _code(b) == codec.FMT_E_SECTION_OUT_OF_BOUNDS


def test_rejects_degenerate_header_fields():
    b = bytearray(29)
    b[0:4] = b"MARS"
    b[20:24] = struct.pack("<f", 8.0)
    b[12], b[13], b[11] = 0, 4, 4
    b[15:17] = struct.pack("<H", 16)
    b[17:19] = struct.pack("<H", 16)
    assert _code(b) == codec.FMT_E_DEGENERATE_HEADER
    # min_size 1 on a 65535 x 65535 header: past the 2^24 leaf-position bound
    c = bytearray(_small_stream())
    c[9] = 1
    c[15:19] = struct.pack("<HH", 65535, 65535)
    assert _code(c) == codec.FMT_E_DEGENERATE_HEADER


def test_rejects_version_qstep_tree_and_payload_errors():
    data = _small_stream()
    b = bytearray(data)
    b[4] = 1
    assert _code(b) == codec.FMT_E_UNSUPPORTED_VERSION
    for q in (float("nan"), float("inf"), 0.5, 65536.0):
        b = bytearray(data)
        b[20:24] = struct.pack("<f", q)
        assert _code(b) == codec.FMT_E_INVALID_QSTEP
    b = bytearray(data)
    b[19] = 0  # no sections
    assert _code(b) == codec.FMT_E_MISSING_TREE
    b = bytearray(data[:29])
    b[25:29] = struct.pack("<I", 4)
    assert _code(bytes(b) + b"\0\0\0\0") == codec.FMT_E_INVALID_TREE_PAYLOAD
    b = bytearray(data)
    b[13] = 21  # bits_beta: an alphabet above the reader's 2^20
    assert _code(b) == codec.FMT_E_UNSUPPORTED


def test_rejects_a_domain_outside_the_plane():
    hdr, leaves = codec.read_mars(_small_stream())
    i = next(k for k, l in enumerate(leaves) if l[3] >= 2)
    l = list(leaves[i])
    l[7] = hdr["height"] - l[2]  # its 2*size window now runs past the bottom
    leaves[i] = tuple(l)
    assert _code(codec.write_mars(hdr, leaves)) == codec.FMT_E_DOMAIN_OUT_OF_BOUNDS
