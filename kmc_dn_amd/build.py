"""Build recipe for libkmc_dn_amd.so (hipcc, gfx950, in-tree).

    python -m kmc_dn_amd.build          # or kmc_dn_amd.build.build()

Each .hip source is compiled to its own object and the objects are linked into
one shared library next to this file.  The FP64 dense kernel is compiled with
-ffp-contract=off so every a*b+c stays a rounded multiply and a rounded add,
as the reference's numba loop evaluates it.  All device code is built with
packed-FP32 instructions disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libkmc_dn_amd.so")
ARCH = os.environ.get("KMC_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
    f"-I{os.path.join(os.path.dirname(HERE), 'include')}",
]
SOURCES = {
    "kmc_plain.hip": ["-ffp-contract=off"],
    "kmc_batch.hip": [],
    "kmc_team.hip": [],
    "kmc_meanfield.hip": ["-ffp-contract=off"],
    "kmc_capi.hip": [],
    "kmc_probe.hip": [],
    "kmc_goshim.cpp": [],
}


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    cc = hipcc()
    with tempfile.TemporaryDirectory(prefix="kmc_build_") as tmp:
        cmds, objs = [], []
        for src, extra in SOURCES.items():
            obj = os.path.join(tmp, os.path.splitext(src)[0] + ".o")
            cmds.append([cc, *COMMON, *extra, "-c", os.path.join(CSRC, src), "-o", obj])
            objs.append(obj)

        def compile_one(cmd):
            if verbose:
                print(" ".join(cmd), flush=True)
            subprocess.run(cmd, check=True)

        jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", "4"))))
        with ThreadPoolExecutor(jobs) as ex:
            list(ex.map(compile_one, cmds))
        tmp_lib = LIB + ".tmp"
        cmd = [cc, *COMMON, "-shared", *objs, "-o", tmp_lib]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        os.replace(tmp_lib, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
