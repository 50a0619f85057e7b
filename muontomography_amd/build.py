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

HIP_SOURCES = ["ctx.hip", "scan.hip", "siddon.hip", "poca.hip", "em.hip", "em32.hip", "map_op.hip", "batch.hip",
               "records.hip"]
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


JSON_DIRS = ["/opt/conda/include", "/usr/include", "/usr/local/include"]
CXX_LIB = os.path.join(HERE, "libmuto_cxx.so")
CLI = os.path.join(HERE, "muto_cli")


def _json_dir():
    for d in JSON_DIRS:
        if os.path.exists(os.path.join(d, "json.hpp")):
            return d
        if os.path.exists(os.path.join(d, "nlohmann", "json.hpp")):
            return os.path.join(d, "nlohmann")
    return None


def build_cxx():
    """The reference-compatible C++ classes (MutoPoCA, MutoMLSD, MutoMAP,
    MutoFile) over libmuto, and the muto_cli driver.  Needs the single-header
    nlohmann json.hpp; without it the C++ layer is skipped with a message."""
    jd = _json_dir()
    if jd is None:
        print("json.hpp not found: the C++ drop-in layer (libmuto_cxx.so, muto_cli) is not built",
              file=sys.stderr)
        return None
    inc = ["-I", os.path.join(ROOT, "include", "muto_cxx"), "-I", os.path.join(ROOT, "include"),
           "-I", jd]
    link = ["-L", HERE, "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link,/opt/rocm/lib"]
    src = os.path.join(HERE, "cxx")
    tmp = CXX_LIB + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", os.path.join(src, "MutoCxx.cpp"),
          "-o", tmp] + inc + link + ["-lmuto"])
    os.replace(tmp, CXX_LIB)
    _run(["g++", "-O2", "-std=c++17", "-Wall", os.path.join(src, "muto_cli.cpp"), "-o", CLI + ".tmp"]
         + inc + link + ["-lmuto_cxx", "-lmuto"])
    os.replace(CLI + ".tmp", CLI)
    # the C++ layer's own check program (tests/cxx), linked like a user's
    chk = os.path.join(ROOT, "tests", "cxx", "dropin_check")
    _run(["g++", "-O2", "-std=c++17", "-Wall", os.path.join(ROOT, "tests", "cxx", "dropin_check.cpp"),
          "-o", chk + ".tmp"] + inc +
         ["-L", HERE, "-Wl,-rpath,$ORIGIN/../../muontomography_amd", "-Wl,-rpath-link,/opt/rocm/lib",
          "-lmuto_cxx", "-lmuto"])
    os.replace(chk + ".tmp", chk)
    return CXX_LIB


def build_oracle():
    """Compile the CPU oracle (test infrastructure) with its Makefile."""
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    return os.path.join(ROOT, "oracle", "liboracle_muto.so")


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_cxx())
    print(build_oracle())
