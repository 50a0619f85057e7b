"""Build libmuto.so (HIP kernels + C ABI, gfx950) in-tree with hipcc.

Usage: python -m muontomography_amd.build   (or build() from Python)
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmuto.so")

HIP_SOURCES = ["ctx.hip", "scan.hip", "siddon.hip", "poca.hip", "em.hip", "em32.hip", "map_op.hip"]
CXX_SOURCES = ["gen.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = [
    "-O3", "-fPIC", "-std=c++17",
    # the reference is FP64 without FMA; every kernel here keeps contraction
    # off and uses fma() explicitly where wanted
    "-ffp-contract=off",
    "-I", os.path.join(ROOT, "include"),
]
HIP_FLAGS = COMMON + [
    "--offload-arch=" + ARCH,
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed: " + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def _obj(src):
    return os.path.join(OBJ, os.path.splitext(src)[0] + ".o")


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    jobs = []
    for s in HIP_SOURCES:
        jobs.append([HIPCC, "-x", "hip", "-c", os.path.join(CSRC, s), "-o", _obj(s)] + HIP_FLAGS)
    for s in CXX_SOURCES:
        jobs.append(["g++", "-c", os.path.join(CSRC, s), "-o", _obj(s), "-fopenmp"] + COMMON)
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        outs = list(ex.map(_run, jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    objs = [_obj(s) for s in HIP_SOURCES + CXX_SOURCES]
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs + ["-fopenmp"])
    os.replace(tmp, LIB)
    return LIB


def build_oracle():
    """Compile the CPU oracle (test infrastructure) with its Makefile."""
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    return os.path.join(ROOT, "oracle", "liboracle_muto.so")


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_oracle())
