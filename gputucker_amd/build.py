"""Builds the gfx950 engine in-tree.

    python -m gputucker_amd.build            # library + driver
    python -m gputucker_amd.build --lib-only

Outputs (git-ignored, shipped with the tree):
    gputucker_amd/libgputucker_hip.so   C ABI of include/gputucker_hip.h
    gputucker_amd/bin/GPUTucker         drop-in CLI driver (-i -o -r -g)
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
DRIVER = os.path.join(HERE, "driver")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgputucker_hip.so")
BIN = os.path.join(HERE, "bin", "GPUTucker")

ARCH = os.environ.get("GT_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -packed-fp32-ops off: packed FP32 VALU forms are not used on this hardware.
# -amdgpu-mfma-vgpr-form: MFMA results land in VGPRs; the delta kernels feed
# every accumulator element to VALU code, which would otherwise cost one
# v_accvgpr_read per element.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-mllvm", "-amdgpu-mfma-vgpr-form",
    "-Wno-unused-command-line-argument", "-Wno-unknown-warning-option",
    "-I", os.path.join(ROOT, "include"),
]

SOURCES = ["gt_sort.hip", "gt_partition.hip", "gt_gram.hip", "gt_gram_mfma.hip", "gt_gram_wide.hip",
           "gt_solve.hip", "gt_abi.hip"]


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr[-6000:]}")
    return obj


def build_lib(jobs: int = 8) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
    os.replace(tmp, LIB)
    return LIB


def build_driver() -> str | None:
    srcs = [os.path.join(DRIVER, f) for f in sorted(os.listdir(DRIVER)) if f.endswith(".cpp")] \
        if os.path.isdir(DRIVER) else []
    if not srcs:
        return None
    os.makedirs(os.path.dirname(BIN), exist_ok=True)
    cmd = ["g++", "-O2", "-std=c++17", "-fopenmp", "-I", os.path.join(ROOT, "include"),
           "-I", DRIVER, *srcs, "-o", BIN, f"-L{HERE}", "-lgputucker_hip",
           "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"driver build failed:\n{r.stderr[-6000:]}")
    return BIN


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib-only", action="store_true")
    ap.add_argument("--jobs", type=int, default=8)
    a = ap.parse_args(argv)
    print("built", build_lib(a.jobs))
    if not a.lib_only:
        b = build_driver()
        if b:
            print("built", b)
    return 0


if __name__ == "__main__":
    sys.exit(main())
