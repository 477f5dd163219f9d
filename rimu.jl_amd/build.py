"""Build ``librimu_hip.so`` (the C-ABI library of ``include/rimu_hip.h``) for gfx950.

Usage: ``python rimu.jl_amd/build.py`` (or through ``__graft_entry__.build()``).
The library is built in-tree next to this file; hipcc cross-compiles without a GPU.
"""
from __future__ import annotations

import os
import pathlib
import subprocess
import sys

HERE = pathlib.Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIB = HERE / "librimu_hip.so"
SOURCES = [CSRC / "rimu_engine.hip", CSRC / "dense_hv.hip", CSRC / "batched.hip"]
DEPENDS = SOURCES + sorted(CSRC.glob("*.hpp")) + [HERE.parent / "include" / "rimu_hip.h"]

# -ffp-contract=off: a*b+c stays two rounded operations (bit-reproducible against the
# oracle, and the same association Julia uses).  Packed FP32 instructions are disabled.
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    return "hipcc"


def build(verbose: bool = False) -> pathlib.Path:
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), *HIPCC_FLAGS, "-o", str(tmp), *map(str, SOURCES)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed ({res.returncode})")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
