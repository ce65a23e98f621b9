"""Builds libfornos_hip.so (gfx950 kernels + host C++) in-tree.

    python -m fornos_amd.build            # or fornos_amd.build.build()

Host C++ is compiled with ROCm's clang++, the kernels with hipcc for gfx950.
Both use -ffp-contract=off so that host-side precomputation and device
arithmetic follow the same operation order without FMA contraction.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libfornos_hip.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
CLANGXX = os.path.join(ROCM, "llvm", "bin", "clang++")

HOST_SOURCES = ["mesh.cpp", "texels.cpp", "bvh.cpp", "fornos.cpp", "imageio.cpp", "meshio.cpp", "runner.cpp"]
HIP_SOURCES = ["kernels.hip", "denoise.hip"]

COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
          "-I", os.path.join(ROOT, "include"), "-Wall", "-Wno-unused-function"]
HIP_FLAGS = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-munsafe-fp-atomics"]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError("build step failed:\n%s\n%s" % (" ".join(cmd), p.stdout))
    return p.stdout


def build(verbose=False):
    """Compiles every source and links the shared library; returns its path."""
    os.makedirs(OBJ, exist_ok=True)
    jobs = []
    for s in HOST_SOURCES:
        o = os.path.join(OBJ, s + ".o")
        jobs.append([CLANGXX, *COMMON, "-c", os.path.join(CSRC, s), "-o", o])
    for s in HIP_SOURCES:
        o = os.path.join(OBJ, s + ".o")
        jobs.append([HIPCC, *COMMON, *HIP_FLAGS, "-c", os.path.join(CSRC, s), "-o", o])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        outs = list(ex.map(_run, jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    objs = [os.path.join(OBJ, s + ".o") for s in HOST_SOURCES + HIP_SOURCES]
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp, "-lpthread", "-lz"])
    os.replace(tmp, LIB)
    return LIB


SHIM_SRC = os.path.join(ROOT, "tests", "cxx", "shim_bake.cpp")
SHIM_BIN = os.path.join(ROOT, "tests", "cxx", "shim_bake")


def build_shim_check():
    """Compiles the C++ host program of the reference-signature classes
    (fornos_amd/csrc/fornos.hpp) that the GPU tests run against the C ABI:
    only the public headers and the built library, as a C++ caller would."""
    tmp = SHIM_BIN + ".tmp"
    _run([CLANGXX, "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), SHIM_SRC, "-o", tmp,
          "-L", HERE, "-lfornos_hip", "-Wl,-rpath,$ORIGIN/../../fornos_amd"])
    os.replace(tmp, SHIM_BIN)
    return SHIM_BIN


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_shim_check())
