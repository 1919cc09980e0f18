"""Build libqandle_sv.so for gfx950 with hipcc (in-tree).

The library links hipRTC for the per-pass specialised kernels it compiles at
run time (qs_jit_*); the kernels shipped in the library itself are built here.

    python -m qandle_amd.csrc.build
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.normpath(os.path.join(HERE, "..", ".."))
OUT_DIR = os.path.normpath(os.path.join(HERE, "..", "_lib"))
OUT = os.path.join(OUT_DIR, "libqandle_sv.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "--offload-arch=gfx950",
    # no packed-FP32 (v_pk_*_f32) code generation
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # keep SimplifyCFG from sinking the per-case register updates of the op
    # dispatch switches into a common block: that turns the register-array
    # index into a phi and forces the whole array into scratch memory
    "-mllvm", "-simplifycfg-sink-common=false",
    # matrix-core (MFMA) results stay in vector registers: no copies between
    # accumulator and vector registers around every fused unitary
    "-mllvm", "-amdgpu-mfma-vgpr-form",
    "-Wno-unused-command-line-argument",
]


def sources():
    return sorted(glob.glob(os.path.join(HERE, "*.hip")))


def build(verbose: bool = False) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    objs = []
    for src in sources():
        obj = os.path.join(OUT_DIR, os.path.basename(src) + ".o")
        cmd = [HIPCC, *FLAGS, "-I", os.path.join(ROOT, "include"), "-I", HERE, "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp,
           "-L/opt/rocm/lib", "-lhiprtc", "-Wl,-rpath,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    for o in objs:
        os.remove(o)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
