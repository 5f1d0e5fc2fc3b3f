"""Build libgkhip.so (all HIP sources under csrc/) for gfx950 with hipcc.

Usage: ``python -m gempy_amd.build`` (or ``__graft_entry__.build()``).
Packed FP32 instructions are disabled (``-packed-fp32-ops``) for every device
compilation; see the package's design notes on the VALU roofline this implies.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libgkhip.so"
INCLUDE = PKG.parent / "include"

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-ffp-contract=fast", "-munsafe-fp-atomics",
    "-Wno-unused-result",
]


def hipcc_path() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build libgkhip.so)")


def sources():
    return sorted(str(p) for p in CSRC.glob("*.hip"))


def build(verbose: bool = False, extra_flags=None) -> Path:
    """Compile every csrc/*.hip into one shared library (always recompiles)."""
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc_path(), *HIPCC_FLAGS, *(extra_flags or []), f"-I{INCLUDE}", "-o", str(tmp), *sources()]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed with exit code {res.returncode}")
    if verbose and res.stderr.strip():
        sys.stderr.write(res.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
