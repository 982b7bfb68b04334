"""Builds mcrt_amd/libmcrt_hip.so in-tree: host C++ with g++ (same fp32
semantics as the reference: -O3, no -march, -ffp-contract=off), HIP sources
with hipcc for gfx950 (no contraction, correctly rounded fp32 division and
square root, denormals kept, packed-fp32 instructions disabled).

    python -m mcrt_amd.build            # or mcrt_amd.build.build()
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_obj")
LIB = os.path.join(HERE, "libmcrt_hip.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = os.environ.get("MCRT_ARCH", "gfx950")

HOST_SRCS = ["capi.cpp", "scene.cpp", "sbvh.cpp", "wide.cpp", "film.cpp"]
HIP_SRCS = ["traverse.hip", "pt.hip"]
DEMOS = ["demo_pt"]                 # C++ programs on the facade include/mcrt.hpp
BIN = os.path.join(HERE, "bin")

INC = ["-I" + os.path.join(ROOT, "include"), "-I" + CSRC, "-I" + os.path.join(ROCM, "include")]
HOST_FLAGS = ["-std=c++17", "-O3", "-fPIC", "-fopenmp", "-ffp-contract=off", "-fvisibility=hidden",
              "-D__HIP_PLATFORM_AMD__", "-Wall", "-Wno-unused-function"]
HIP_FLAGS = ["-std=c++17", "-O3", "-fPIC", "--offload-arch=" + ARCH, "-ffp-contract=off",
             "-fno-fast-math", "-fhip-fp32-correctly-rounded-divide-sqrt",
             "-fno-gpu-flush-denormals-to-zero", "-fvisibility=hidden",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build(verbose=False, jobs=8):
    """Compiles every source (always, even when the library exists)."""
    os.makedirs(OBJ, exist_ok=True)
    hipcc = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
    steps = []
    for s in HOST_SRCS:
        o = os.path.join(OBJ, s + ".o")
        steps.append(["g++"] + HOST_FLAGS + INC + ["-c", os.path.join(CSRC, s), "-o", o])
    for s in HIP_SRCS:
        o = os.path.join(OBJ, s + ".o")
        steps.append([hipcc] + HIP_FLAGS + INC + ["-c", os.path.join(CSRC, s), "-o", o])
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        outs = list(ex.map(_run, steps))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    objs = [os.path.join(OBJ, s + ".o") for s in HOST_SRCS + HIP_SRCS]
    tmp = LIB + ".tmp"
    # g++ links libgomp (the OpenMP runtime torch also uses); hipcc -fopenmp
    # would pull in LLVM libomp and two OpenMP runtimes would share the process.
    _run(["g++", "-shared", "-fopenmp", "-o", tmp] + objs + ["-L" + os.path.join(ROCM, "lib"), "-lamdhip64",
                                                             "-Wl,-rpath," + os.path.join(ROCM, "lib")])
    os.replace(tmp, LIB)
    os.makedirs(BIN, exist_ok=True)
    for d in DEMOS:
        _run(["g++", "-std=c++17", "-O2", "-Wall", "-I" + os.path.join(ROOT, "include"),
              os.path.join(HERE, "demo", d + ".cpp"), "-o", os.path.join(BIN, d), "-L" + HERE, "-lmcrt_hip",
              "-Wl,-rpath,$ORIGIN/.."])
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
