"""Build libgrains_hip.so for gfx950 with hipcc (in-tree, next to this file).

FP64 translation units are compiled with -ffp-contract=off so every product
and sum rounds separately, as in the reference CPU code; the FP32-mixed unit
allows contraction.  Packed FP32 instructions are disabled for all device
code.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgrains_hip.so")
ARCH = os.environ.get("GRAINS_OFFLOAD_ARCH", "gfx950")

UNITS = {
    "gr_broad": "off",
    "gr_narrow64": "off",
    "gr_narrow32": "fast",
    "gr_capi": "off",
}
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Xclang", "-target-feature",
          "-Xclang", "-packed-fp32-ops", "-Wno-unused-result"]


def _compile(unit, contract, hipcc):
    src = os.path.join(CSRC, unit + ".hip")
    obj = os.path.join(BUILD, unit + ".o")
    cmd = [hipcc] + COMMON + [f"-ffp-contract={contract}", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {unit}:\n{r.stderr[-4000:]}")
    return obj


def build(verbose=False):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(UNITS)) as ex:
        objs = list(ex.map(lambda kv: _compile(kv[0], kv[1], hipcc), UNITS.items()))
    tmp = LIB + ".tmp"
    cmd = [hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
