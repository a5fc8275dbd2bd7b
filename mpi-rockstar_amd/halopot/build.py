"""Builds libhalopot.so (gfx950) in-tree with hipcc.

One translation unit, csrc/hp_lib.hip, which includes the kernel files.
Flags:
  -ffp-contract=off   expressions that must round like the reference's C99
                      code are never fused; hot loops use explicit fmaf
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed FP32 instructions are generated
  -fvisibility=hidden only the C ABI of include/halopot.h is exported
"""
from __future__ import annotations

import os
import pathlib
import shutil
import subprocess
import sys

PKG = pathlib.Path(__file__).resolve().parent
ROOT = PKG.parent.parent
CSRC = PKG.parent / "csrc"
LIB = PKG / "libhalopot.so"


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and pathlib.Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: libhalopot needs the ROCm toolchain")


def command(out: pathlib.Path = LIB, extra=()) -> list[str]:
    return [
        hipcc(), "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
        "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-fvisibility=hidden", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
        f"-I{ROOT / 'include'}", *extra, "-o", str(out), str(CSRC / "hp_lib.hip"),
    ]


def build(verbose: bool = False) -> pathlib.Path:
    tmp = LIB.with_suffix(".so.tmp")
    cmd = command(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + res.stdout + res.stderr)
    if verbose and res.stderr.strip():
        print(res.stderr, file=sys.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
