"""Builds the luce-render HIP library (liblrhip.so) in-tree for gfx950.

    python luce-render_amd/build.py

Every HIP and C++ source under csrc/ goes through one hipcc link. Flags:
-ffp-contract=off keeps f32 arithmetic in the reference's rounding (no fused
multiply-add), and packed-fp32 instructions are disabled.
"""
import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OUT = HERE / "luce_render" / "liblrhip.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-pthread",
    "-munsafe-fp-atomics",
]


def sources():
    return sorted(str(p) for p in CSRC.iterdir() if p.suffix in (".hip", ".cpp"))


def build(verbose=False):
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [HIPCC, *FLAGS, "-o", str(tmp), *sources()]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-q" not in sys.argv))
