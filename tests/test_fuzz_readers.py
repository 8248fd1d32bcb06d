"""The stream readers under a sanitizer fuzz loop (CPU).

tests/fuzz/fuzz_read.cpp, built by tests/fuzz/Makefile with AddressSanitizer
and UndefinedBehaviorSanitizer over the library's own codec.cpp, mutates seed
streams (real `.mars` encodes of this chain and golden `.ifs` fixtures) and
feeds every mutant to mars_read and mars_ifs_parse, decoding one iteration of
whatever they accept -- the reference's fuzz target
(crates/mars-codec/fuzz/fuzz_targets/mars_format_read.rs:1-17) with a
mutation loop in place of libFuzzer.  A sanitizer report, a crash or a
positive status fails the run; every typed error the readers have must occur.
"""

import shutil
import subprocess
from pathlib import Path

import pytest

from mars_amd import synth
from mars_amd.codec import EncodeOptions, encode_plane
from mars_amd.provider import EncodeParams
from tests import _oracle as O

FUZZ = Path(__file__).resolve().parent / "fuzz"
GOLDEN = Path(__file__).resolve().parent / "golden" / "mars1"


@pytest.mark.timeout(600)
def test_readers_survive_mutated_streams(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("no g++ to build the sanitizer driver")
    subprocess.run(["make", "-s", "-C", str(FUZZ)], check=True)
    seeds = []
    for i, (img, p) in enumerate([
        (synth.plane(64, 48, 1), EncodeParams(min_size=4, max_size=16, shift=4)),
        (synth.plane(48, 40, 5), EncodeParams(min_size=4, max_size=16, shift=4, lambda_=40.0)),
        (synth.plane(37, 29, 6), EncodeParams(min_size=2, max_size=8, shift=2, lambda_=100.0)),
    ]):
        prov = O.Provider(img)
        data, _ = encode_plane(img, p, EncodeOptions(), prov.fn, prov.user)
        path = tmp_path / f"seed{i}.mars"
        path.write_bytes(data)
        seeds.append(str(path))
    seeds += [str(GOLDEN / "tiny64__walkthrough__fisher__r8.ifs"),
              str(GOLDEN / "mixed_129x127__min2__fisher__r8.ifs")]
    out = subprocess.run([str(FUZZ / "_build" / "fuzz_read"), "60000", "7", *seeds],
                         capture_output=True, text=True, timeout=500)
    print(out.stdout)
    assert out.returncode == 0, out.stderr[-3000:]
    lines = dict(l.split(":", 1) for l in out.stdout.splitlines() if ":" in l)
    codes = {int(kv.split(":")[0]) for kv in lines["mars_read"].split()}
    # every MarsFormatError variant this reader has, and accepted streams
    assert codes >= {0, -101, -102, -103, -104, -105, -106, -107, -108, -109, -112}, codes
    ifs_codes = {int(kv.split(":")[0]) for kv in lines["ifs_parse"].split()}
    assert ifs_codes >= {0, -104, -105, -109, -110, -111}, ifs_codes
