"""CPU checks of the C ABI library: it loads and exports what include/mars_hip.h declares."""

import ctypes
import re
from pathlib import Path

import pytest

from mars_amd import _native as N

ROOT = Path(__file__).resolve().parent.parent


def _declared(header: str):
    text = (ROOT / "include" / header).read_text()
    return sorted(set(re.findall(r"\b(mars_hip_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    if not N.LIB_PATH.exists():
        pytest.fail("libmars_hip.so is missing: __graft_entry__.build() must run first")
    lib = ctypes.CDLL(str(N.LIB_PATH))
    declared = _declared("mars_hip.h")
    assert declared, "no declarations parsed"
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing
    assert set(declared) == set(N.EXPORTS)
    codec = [s for s in re.findall(r"\b(mars_[a-z0-9_]+)\s*\(",
                                   (ROOT / "include" / "mars_codec.h").read_text())]
    assert codec and all(hasattr(lib, s) for s in codec), codec


def test_null_arguments_are_rejected_without_a_device():
    L = N.lib()
    assert L.mars_hip_open(0, None) == N.E_ARG
    assert L.mars_hip_last_error(None) == b"null context"
