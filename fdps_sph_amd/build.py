"""Builds libfdps_sph_gpu.so (gfx950) in-tree with hipcc.

Each HIP source is compiled to an object in csrc/_build/, then linked into
fdps_sph_amd/libfdps_sph_gpu.so.  The FP64 path (interaction kernels, tree,
EoS, integrator) is compiled with -ffp-contract=off so that every expression
is evaluated operation by operation, as the reference's FP64 arithmetic; the
mixed-precision kernels allow contraction.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(CSRC, "_build")
LIB = os.path.join(HERE, "libfdps_sph_gpu.so")
ARCH = os.environ.get("FDPS_SPH_ARCH", "gfx950")

COMMON = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-Wall", "-Wno-unused-function",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
# The support masks run on the matrix cores; their accumulators stay in
# ordinary VGPRs (gfx950 allows it) instead of a separate block of AGPRs that
# would cost waves of occupancy.
MFMA_VGPR = ["-mllvm", "-amdgpu-mfma-vgpr-form"]
SOURCES = {
    "kernels_fp64.hip": ["-ffp-contract=off", *MFMA_VGPR],
    "tree.hip": ["-ffp-contract=off", *MFMA_VGPR],
    "sort.hip": [],
    "eos.hip": ["-ffp-contract=off"],
    "integrate.hip": ["-ffp-contract=off"],
    "api.hip": ["-ffp-contract=off"],
    "kernels_mixed.hip": ["-ffp-contract=fast", *MFMA_VGPR],
}


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src: str, extra: list[str]) -> str:
    obj = os.path.join(BUILD, src.replace(".hip", ".o"))
    cmd = [hipcc(), *COMMON, *extra, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    """Compile every HIP source for gfx950 and link the shared library."""
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        futs = {ex.submit(_compile, s, e): s for s, e in SOURCES.items()}
        objs = []
        for f in cf.as_completed(futs):
            objs.append(f.result())
            if verbose:
                print("compiled", futs[f], flush=True)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *sorted(objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB, flush=True)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
