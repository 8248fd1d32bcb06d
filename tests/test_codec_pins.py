"""The encoder chain above the search, pinned to the reference's own stream hashes.

crates/mars-codec/tests/search_provider.rs:13-67 and 179-227: four encodes of a
closed-form 16x16 plane whose `.mars` bytes (length + SHA-256) and search eval
counts the reference pins; and the CLI-A stream (crates/mars-cli/tests/
cli_a_gate.rs:55-56, 224-249), a 128x128 RD encode with all four leaf modes in
a MARC grey container.  The chain here (walks, RD, rate snapshot, format,
rANS) runs with the oracle's exhaustive search as its provider, so these pin
both the oracle search and the host-side chain.  CPU only.
"""

import hashlib

import numpy as np
import pytest

from mars_amd.codec import EncodeOptions, encode_plane, wrap_gray
from mars_amd.provider import EncodeParams
from tests import _oracle as O


def pinned_image() -> np.ndarray:
    i = np.arange(256)
    v = np.where(i % 16 < 8, 100, (i * 37 + i // 16 * 13) % 256)
    return v.astype(np.uint8).reshape(16, 16)


def params(lam):
    return EncodeParams(min_size=4, max_size=8, shift=4, bits_alfa=4, bits_beta=7, max_alfa=1.0,
                        t_rms=8.0, zero_threshold=0, lambda_=lam)


CASES = [
    (params(None), EncodeOptions(), 53, 608,
     "bb480414035f577917542c7f972cb302f904cecbc79449568d05ad74345e98d8"),
    (params(2.0), EncodeOptions(), 149, 1184,
     "5069a669a65b6926bc8a597099e0c40c36ac94f038138480226c9b0385b6ae1c"),
    (params(2.0), EncodeOptions(adaptive_density=True, lambda_adaptive_qstep=True), 161, 864,
     "afbb23207c8e98b70a1b23639701a8c02d46bd10e7299c58cebe60d52cee76d8"),
    (params(50.0), EncodeOptions(allowed_modes=(True, False, True, False),
                                 fixed_qstep=17.123456789), 53, 1184,
     "44fd4f749006cddd0054b817ba5afd630f356bfa6c18f5b68473e4a300c3885f"),
]


@pytest.mark.parametrize("case", range(4))
def test_pinned_pre_provider_streams(case):
    p, o, length, evals, digest = CASES[case]
    img = pinned_image()
    prov = O.Provider(img)
    data, st = encode_plane(img, p, o, prov.fn, prov.user)
    assert st.search_evals == evals
    assert st.warmup_evals == (608 if p.lambda_ is not None else 0)
    assert len(data) == length
    assert hashlib.sha256(data).hexdigest() == digest


def test_mars_header_layout():
    img = pinned_image()
    prov = O.Provider(img)
    data, _ = encode_plane(img, params(None), EncodeOptions(), prov.fn, prov.user)
    assert data[:4] == b"MARS" and data[4] == 0 and data[6] == 1 and data[7] == 8
    assert (data[9], data[10], data[11], data[12], data[13], data[14]) == (4, 8, 4, 4, 7, 32)
    assert int.from_bytes(data[15:17], "little") == 16
    assert int.from_bytes(data[17:19], "little") == 16
    assert data[19] == 1 and data[24] == 1
    assert int.from_bytes(data[25:29], "little") == len(data) - 29
    wrapped = wrap_gray(data)
    assert wrapped[:4] == b"MARC" and wrapped[4:7] == bytes([0, 0, 1])
    assert int.from_bytes(wrapped[7:11], "little") == len(data) and wrapped[11:] == data


# crates/mars-cli/tests/cli_a_gate.rs:55-56, 224-249: `encmars` on a 128x128
# gentle gradient (100 + (r + c) % 5), lambda 200, modes 0-3, t_rms 0 and the
# CLI defaults (sizes 4..16, shift 4, bits 4/7, max_alfa 1, fixed residual
# step 8, rd_candidates 1) writes a MARC grey container with this SHA-256 when
# --adaptive-density is omitted; with it the stream must differ.
CLI_A_STREAM_SHA256 = "55ccaacf9b9c0e952673992404cfdf8ebb21e8503ae3d45a67b4e82a5eed3544"


def cli_a_gradient() -> np.ndarray:
    r, c = np.mgrid[0:128, 0:128]
    return (100 + (r + c) % 5).astype(np.uint8)


def cli_a_params():
    return EncodeParams(min_size=4, max_size=16, shift=4, bits_alfa=4, bits_beta=7, max_alfa=1.0,
                        t_rms=0.0, zero_threshold=0, lambda_=200.0)


def test_cli_a_default_stream():
    img = cli_a_gradient()
    streams = []
    for density in (False, True):
        prov = O.Provider(img)
        data, _ = encode_plane(img, cli_a_params(),
                               EncodeOptions(allowed_modes=(True, True, True, True),
                                             adaptive_density=density), prov.fn, prov.user)
        streams.append(wrap_gray(data))
    assert hashlib.sha256(streams[0]).hexdigest() == CLI_A_STREAM_SHA256
    assert streams[1] != streams[0]
