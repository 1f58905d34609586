"""Build libs2hip.so (HIP kernels for gfx950 + C ABI) in-tree.

Each translation unit is compiled to an object with its own flags, in parallel,
then linked with hipcc.  Exactness-critical units (voxelization, exact counts and
the curve fold) are compiled with -ffp-contract=off so no FMA changes an f64
result; every HIP unit disables packed FP32 instructions.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIB = HERE / "libs2hip.so"
OBJ = HERE / "_obj"

ARCH = os.environ.get("S2HIP_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wno-unused-command-line-argument"]
UNITS = {
    "s2hip_runtime.hip": ["-ffp-contract=off"],
    "s2hip_exact.hip": ["-ffp-contract=off"],
    "s2hip_voxel.hip": ["-ffp-contract=off"],
    "s2hip_mc.hip": ["-ffp-contract=off"],
    "s2hip_fft.hip": ["-ffp-contract=fast"],
    "s2hip_batch.hip": ["-ffp-contract=off"],
    "s2hip_serve.hip": ["-ffp-contract=off"],
    "s2hip_probe.hip": [],
    "s2hip_meshmc.hip": ["-ffp-contract=off"],
    "s2hip_volin.hip": ["-ffp-contract=off"],
    "s2hip_meshin.hip": ["-ffp-contract=off"],
}
HOST_UNITS = {"s2hip_host.cpp": ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off"],
              "s2hip_volio.cpp": ["-O2", "-std=c++17", "-fPIC"],
              "s2hip_volfrac.cpp": ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off"]}


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return r


def build(verbose: bool = False) -> Path:
    OBJ.mkdir(exist_ok=True)
    jobs = []
    for src, flags in UNITS.items():
        obj = OBJ / (src + ".o")
        jobs.append([HIPCC, *COMMON, *flags, "-c", str(CSRC / src), "-o", str(obj)])
    for src, flags in HOST_UNITS.items():
        obj = OBJ / (src + ".o")
        jobs.append(["g++", *flags, "-c", str(CSRC / src), "-o", str(obj)])
    workers = min(len(jobs), max(1, (os.cpu_count() or 2)))
    with ThreadPoolExecutor(max_workers=workers) as ex:
        for r in ex.map(_run, jobs):
            if verbose and (r.stdout or r.stderr):
                print(r.stdout, r.stderr, file=sys.stderr)
    objs = [str(OBJ / (s + ".o")) for s in list(UNITS) + list(HOST_UNITS)]
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-lpthread", "-lz", "-o", str(tmp)])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
