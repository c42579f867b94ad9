"""Builds the HIP library in-tree for gfx950: adgpu/libadgpu_hip.so.

python -m adgpu.build  (from autodock-gpu_amd/), or adgpu.build.build().
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
REPO = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OUT = os.path.join(HERE, "libadgpu_hip.so")

SOURCES = ["adgpu_kernels.hip", "adgpu_host.cpp"]
ARCH = os.environ.get("ADGPU_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def flags():
    return ["-O3", "-std=c++17", "-fPIC", "--offload-arch=%s" % ARCH,
            # no packed FP32 instructions (v_pk_*_f32) in device code
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-I", os.path.join(REPO, "include"), "-Wno-unused-command-line-argument"]


def build(verbose: bool = False) -> str:
    objs = []
    tmp = os.path.join(PKG, "_build")
    os.makedirs(tmp, exist_ok=True)
    for src in SOURCES:
        obj = os.path.join(tmp, src + ".o")
        cmd = [hipcc()] + flags() + ["-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".hip"):
            cmd[1:1] = ["-x", "hip"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp_out = OUT + ".tmp"
    cmd = [hipcc(), "--offload-arch=%s" % ARCH, "-shared", "-fPIC", "-o", tmp_out] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp_out, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
