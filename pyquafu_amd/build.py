"""Build libqfvm_hip.so in-tree for gfx950 with hipcc.

python -m pyquafu_amd.build        # compile every source, link pyquafu_amd/libqfvm_hip.so
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libqfvm_hip.so")
BUILD = os.path.join(CSRC, "_build")
INCLUDE = os.path.abspath(os.path.join(HERE, "..", "include"))

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
COMMON = ["-O3", "-fPIC", "-std=c++17", f"-I{INCLUDE}", f"-I{CSRC}", "-Wall", "-Wno-unused-function"]
# Device code: gfx950 only; packed-FP32 instructions disabled (they have produced wrong results on
# these machines); hipcc's host pass warns that it does not know the feature, which is harmless.
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
# The fused kernel indexes its 16-amplitude register array with compile-time constants in every
# (gate kind, target bit) case; SimplifyCFG's sinking of identical-looking cases would merge them
# into one block with a runtime index and push the array to scratch memory.
DEVICE += ["-mllvm", "-simplifycfg-sink-common=false"]
# The fused kernels dispatch on op records through nested switches whose conditions are uniform
# (scalar registers); structurizing those regions into predicated flow blocks costs register moves
# at every join (measured on the row-of-8 kernel: 112 instead of 122 VGPRs, 30% fewer moves).
DEVICE += ["-mllvm", "-structurizecfg-skip-uniform-regions"]

SOURCES = ["gates.hip", "reduce.hip", "sample.hip", "fused.hip", "plan.cpp", "capi.cpp"]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *COMMON, *DEVICE, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    shutil.move(tmp, OUT)
    if verbose:
        print(f"built {OUT}")
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
