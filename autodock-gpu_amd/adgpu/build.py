"""Builds the HIP library in-tree for gfx950: adgpu/libadgpu_hip.so.

python -m adgpu.build  (from autodock-gpu_amd/), or adgpu.build.build().
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
REPO = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OUT = os.path.join(HERE, "libadgpu_hip.so")

# one translation unit per kernel family, compiled in parallel
SOURCES = ["k_ls_ad.hip", "k_ls_sw.hip", "k_ls_adam.hip", "k_ls_fire.hip", "k_ga.hip", "k_eval.hip", "k_rescore.hip", "k_autostop.hip", "k_calib.hip",
           "adgpu_host.cpp", "adgpu_serve.cpp"]
ARCH = os.environ.get("ADGPU_OFFLOAD_ARCH", "gfx950")

# Instruction scheduling per translation unit, measured on the 30-atom ligand
# (LGA, nrun 100 x pop 150, 2.5 M evaluations per run; M evals/s against the
# default scheduler): Solis-Wets with the AMDGPU register-pressure trackers
# 1280-1285 against 1263-1267, no variant spilling more.  The other local
# searches keep the default: with the trackers ADADELTA ran at 694 against
# 737, FIRE 806 against 827, ADAM 771 against 787; the ILP strategy gave ADAM
# 807 against 787 on this ligand but slowed the served FP16 screen (10-60
# atoms) by 1.6 %.  A scheduler reorders instructions only, so results are
# bit-identical.
SOURCE_FLAGS = {
    "k_ls_sw.hip": ["-mllvm", "-amdgpu-use-amdgpu-trackers"],
}


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


def _compile(src: str, tmp: str, verbose: bool, defines=()) -> str:
    obj = os.path.join(tmp, src + ".o")
    cmd = [hipcc()] + flags() + SOURCE_FLAGS.get(src, []) + ["-D%s" % d for d in defines] + \
        ["-c", os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".hip"):
        cmd[1:1] = ["-x", "hip"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return obj


def build(verbose: bool = False, jobs: int = 0, defines=(), out: str = OUT, tmp: str | None = None) -> str:
    """Compiles every source and links `out`.  defines: extra -D macros
    (e.g. ADGPU_WAVES_PER_EU=5) for a variant build written elsewhere."""
    tmp = tmp or os.path.join(PKG, "_build")
    os.makedirs(tmp, exist_ok=True)
    jobs = jobs or max(1, min(len(SOURCES), int(os.environ.get("MAX_JOBS", "0")) or (os.cpu_count() or 1), 16))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, tmp, verbose, defines), SOURCES))
    tmp_out = out + ".tmp"
    cmd = [hipcc(), "--offload-arch=%s" % ARCH, "-shared", "-fPIC", "-o", tmp_out] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp_out, out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
