"""Build the in-tree native libraries (run from __graft_entry__.build()).

* ``mars_amd/libmars_hip.so`` -- the HIP search backend for gfx950 (hipcc).
  Device code is compiled with ``-ffp-contract=off`` (the binary64 fit must
  round exactly like the reference) and with packed FP32 instructions
  disabled.
"""

from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "mars_amd"
CSRC = PKG / "csrc"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MARS_HIP_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # MFMA accumulators in VGPRs: the epilogue reads every accumulator, so
    # AGPR copies would cost one instruction per value.
    "-mllvm", "-amdgpu-mfma-vgpr-form",
    "-Wall", "-Wno-unused-function",
    f"-I{ROOT / 'include'}", f"-I{CSRC}",
]


def _run(cmd: list[str]) -> None:
    print("+", " ".join(str(c) for c in cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass warns that it does not know the packed-fp32 feature.
    err = "\n".join(l for l in proc.stderr.splitlines() if "packed-fp32-ops" not in l)
    if err.strip():
        print(err, file=sys.stderr)
    if proc.returncode != 0:
        raise RuntimeError(f"build failed ({proc.returncode}): {' '.join(map(str, cmd))}")


CXX = os.environ.get("CXX", "g++")
# The CPU encoder chain: binary64 must round like the reference's Rust
# (no FMA contraction, no fast-math).
CXX_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-Wall",
             f"-I{ROOT / 'include'}", f"-I{CSRC}", "-D__HIP_PLATFORM_AMD__",
             "-I/opt/rocm/include"]


def build_hip() -> Path:
    out = PKG / "libmars_hip.so"
    objs = []
    for src in ("kernels.hip", "fisher.hip", "api.hip"):
        obj = CSRC / (Path(src).stem + ".o")
        _run([HIPCC, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(obj)
    obj = CSRC / "codec.o"
    _run([CXX, *CXX_FLAGS, "-c", str(CSRC / "codec.cpp"), "-o", str(obj)])
    objs.append(obj)
    tmp = out.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, out)
    for o in objs:
        o.unlink()
    return out


def build_all() -> None:
    build_hip()


if __name__ == "__main__":
    build_all()
