"""Build librivrs_ldlt.so (gfx950 HIP kernels + host driver + C ABI) in this directory.

Host-only sources are compiled with g++, HIP sources with hipcc --offload-arch=gfx950; the
shared library is linked by hipcc.  Objects go to _build/ next to this file.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "librivrs_ldlt.so")

HOST_SOURCES = ["symbolic.cpp", "structure.cpp", "matching.cpp"]
HIP_SOURCES = ["numeric.hip", "solve.hip", "refine.hip", "mlp.hip", "capi.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]
HOST_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-Wall"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
    return r


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    for s in HOST_SOURCES:
        obj = os.path.join(BUILD, s + ".o")
        jobs.append(["g++", *HOST_FLAGS, "-c", os.path.join(CSRC, s), "-o", obj])
    for s in HIP_SOURCES:
        obj = os.path.join(BUILD, s + ".o")
        extra = ["-x", "hip"] if s.endswith(".cpp") else []
        jobs.append([HIPCC, *HIP_FLAGS, *extra, "-c", os.path.join(CSRC, s), "-o", obj])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for r in ex.map(_run, jobs):
            if verbose and (r.stdout or r.stderr):
                sys.stderr.write(r.stdout + r.stderr)
    objs = [os.path.join(BUILD, s + ".o") for s in HOST_SOURCES + HIP_SOURCES]
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
