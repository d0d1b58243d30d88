"""Build recipe for libxitils_amd.so (gfx950) and the C++ API example.

Every HIP/C++ file is compiled by hipcc with -ffp-contract=off (one IEEE
rounding per source operation, the float behaviour the CPU checker follows)
and packed-FP32 instructions disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_lib")
LIB = os.path.join(OUT_DIR, "libxitils_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else "hipcc")

SOURCES = ["runtime.hip", "kernels.hip", "guide.hip", "vmf.hip", "bvh.cpp"]
# the guiding network's MFMA accumulators in ordinary VGPRs (the VALU reads
# them for the activations; in AGPRs every element costs an extra copy)
PER_FILE = {"guide.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}
COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} ... {cmd[-1]}")
    return r


def build(verbose: bool = False, out_lib: str = LIB, defines=()) -> str:
    out_dir = os.path.dirname(out_lib)
    os.makedirs(out_dir, exist_ok=True)
    objs = []
    for src in SOURCES:
        obj = os.path.join(out_dir, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC, "--offload-arch=gfx950", *COMMON, *PER_FILE.get(src, []), *[f"-D{d}" for d in defines], "-I",
               os.path.join(ROOT, "include"), "-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd = [HIPCC, *COMMON, *[f"-D{d}" for d in defines], "-x", "c++", "-I", os.path.join(ROOT, "include"), "-c",
                   os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        objs.append(obj)
    tmp = out_lib + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"])
    os.replace(tmp, out_lib)
    for o in objs:
        os.remove(o)
    return out_lib


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
