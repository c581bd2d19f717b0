"""Build liblatec.so in-tree (latec_amd/_lib/) with hipcc for gfx950.

    python -m latec_amd.csrc.build          # or build() from Python

HIP sources are compiled with -ffp-contract=off (the codec's FP64 arithmetic
is specified op by op) and with packed-FP32 instructions disabled.
"""

from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
OUT_DIR = os.path.join(PKG, "_lib")
OBJ_DIR = os.path.join(OUT_DIR, "obj")
LIB = os.path.join(OUT_DIR, "liblatec.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ZSTD_INC = os.environ.get("LATEC_ZSTD_INC", "/opt/conda/include")
ARCH = os.environ.get("LATEC_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["latec_interp.hip", "latec_sheet.hip", "latec_tables.hip", "latec_gnn.hip", "latec_train.hip",
               "latec_mlrans.hip",
               "latec_capi.cpp"]
CXX_SOURCES = ["latec_plan.cpp", "latec_entropy.cpp"]

COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", f"-I{ZSTD_INC}",
          f"-I{os.path.join(os.path.dirname(PKG), 'include')}"]
# MFMA results in VGPRs (not AGPRs): the GNN kernels apply GELU / softmax to
# every accumulator, which an AGPR destination would first copy out with one
# v_accvgpr_read each
HIP_FLAGS = COMMON + [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
                      "-packed-fp32-ops", "-munsafe-fp-atomics",
                      "-mllvm", "-amdgpu-mfma-vgpr-form"]


def _compile(src: str) -> str:
    path = os.path.join(HERE, src)
    obj = os.path.join(OBJ_DIR, os.path.basename(src) + ".o")
    if src in CXX_SOURCES:
        cmd = ["g++", *COMMON, "-c", path, "-o", obj]
    else:
        cmd = [HIPCC, *HIP_FLAGS, "-x", "hip", "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ_DIR, exist_ok=True)
    sources = [s for s in HIP_SOURCES + CXX_SOURCES if os.path.exists(os.path.join(HERE, s))]
    with ThreadPoolExecutor(max_workers=min(8, len(sources))) as ex:
        objs = list(ex.map(_compile, sources))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp,
           "-L/usr/lib/x86_64-linux-gnu", "-l:libzstd.so.1", "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
