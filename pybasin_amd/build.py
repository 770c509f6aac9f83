"""Build libbasinhip.so (HIP, gfx950) in-tree with hipcc.

``python -m pybasin_amd.build`` compiles every HIP source of the package into
``pybasin_amd/libbasinhip.so``.  hipcc cross-compiles for gfx950 without a GPU.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libbasinhip.so")

# translation units and their extra flags: the per-lane solver is scheduled
# for instruction-level parallelism (a wave alone on its SIMD waits on its
# dependent arithmetic; measured 3 % faster on the Duffing launch)
SOURCES = {
    "basinhip_api.hip": [],
    "lane_api.hip": ["-mllvm", "-amdgpu-sched-strategy=max-ilp"],
}

# Packed FP32 (v_pk_*_f32) is disabled for every HIP object: some of its
# forms have produced silently wrong results on these devices.
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found; set HIPCC or install ROCm")


def build(verbose: bool = False) -> str:
    """Compile every translation unit (in parallel) and link the library."""
    tmpdir = LIB + ".objs"
    os.makedirs(tmpdir, exist_ok=True)
    procs, objs = [], []
    for src, extra in SOURCES.items():
        obj = os.path.join(tmpdir, src.replace(".hip", ".o"))
        cmd = [hipcc(), *HIPCC_FLAGS, *extra, "-I", os.path.join(ROOT, "include"), "-c", "-o", obj,
               os.path.join(CSRC, src)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd)))
        objs.append(obj)
    for cmd, pr in procs:
        if pr.wait() != 0:
            raise subprocess.CalledProcessError(pr.returncode, cmd)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    shutil.rmtree(tmpdir, ignore_errors=True)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
