"""Builds the luce-render HIP library (liblrhip.so) in-tree for gfx950.

    python luce-render_amd/build.py

Every HIP and C++ source under csrc/ is compiled by its own hipcc process
(in parallel), then one hipcc link makes the library. Flags:
-ffp-contract=off keeps f32 arithmetic in the reference's rounding (no fused
multiply-add), and packed-fp32 instructions are disabled.
"""
import os
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor
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
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-pthread",
    "-munsafe-fp-atomics",
]


def sources():
    return sorted(str(p) for p in CSRC.iterdir() if p.suffix in (".hip", ".cpp"))


def build(verbose=False, jobs=None):
    tmp = OUT.with_suffix(".so.tmp")
    srcs = sources()
    jobs = jobs or min(len(srcs), max(1, min(16, os.cpu_count() or 1)))
    with tempfile.TemporaryDirectory(prefix="lrhip_") as work:
        objs = [str(Path(work) / (Path(src).name + ".o")) for src in srcs]

        def compile_one(i):
            cmd = [HIPCC, *FLAGS, "-c", "-o", objs[i], srcs[i]]
            if verbose:
                print(" ".join(cmd), flush=True)
            return subprocess.run(cmd, capture_output=True, text=True)

        with ThreadPoolExecutor(jobs) as pool:
            results = list(pool.map(compile_one, range(len(srcs))))
        failed = [(srcs[i], r) for i, r in enumerate(results) if r.returncode != 0]
        for src, r in failed:
            sys.stderr.write(f"--- {src}\n{r.stdout}{r.stderr}")
        if failed:
            raise subprocess.CalledProcessError(failed[0][1].returncode, "hipcc " + failed[0][0])
        cmd = [HIPCC, *FLAGS, "-shared", "-o", str(tmp), *objs]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-q" not in sys.argv))
