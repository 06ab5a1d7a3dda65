"""Builds libnsp.so (HIP kernels for gfx950 + the C ABI) and the CPU checker.

Objects go to math_amd/_build/, the shared library to math_amd/libnsp.so
(in-tree, so it travels with the repository snapshot).  Every call compiles.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libnsp.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = [
    "-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
    # packed FP32 instructions stay off (they are formed by -O3 from float code)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", "-Wno-unknown-pragmas",
]

# per-source extra flags
SOURCES = {
    "nsp_build.hip": ["-ffp-contract=off"],   # FP64 records bit-equal to the CPU checker
    "nsp_render.hip": [],
    "nsp_denoise.hip": [],
    "nsp_raster.hip": [],
    "nsp_shadow_raster.hip": [],
    "nsp_hull.hip": [],
    "nsp_cocone.hip": ["-ffp-contract=off"],  # FP64 reconstruction equal to the CPU checker's
    "nsp_api.cpp": ["-x", "hip"],
    "nsp_meshgen.cpp": ["-x", "hip"],
    "nsp_obj.cpp": ["-ffp-contract=off"],   # float operations as written (file values)
}


def _compile(src, extra, verbose):
    obj = os.path.join(OUT, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON + extra + ["-c", os.path.join(CSRC, src), "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{r.stdout}\n{r.stderr}")
    return obj


def build_library(verbose=False):
    os.makedirs(OUT, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=4) as ex:
        futs = [ex.submit(_compile, s, e, verbose) for s, e in SOURCES.items()]
        objs = [f.result() for f in futs]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


def build_oracle(verbose=False):
    d = os.path.join(ROOT, "oracle")
    r = subprocess.run(["make", "-B", "-C", d], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(r.stdout)
    return os.path.join(d, "libnsp_oracle.so")


def build_cpp_tests(verbose=False):
    """The C++17 drop-in tests (tests/cpp), compiled by the host compiler
    against include/nsp.h and libnsp.so."""
    exes = []
    for name in ("test_painter_api", "test_hull_order", "test_conventions", "test_reconstruction"):
        src = os.path.join(ROOT, "tests", "cpp", name + ".cpp")
        exe = os.path.join(ROOT, "tests", "cpp", name)
        cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"), "-I", CSRC, src, "-o", exe,
               "-L", HERE, "-lnsp", "-Wl,-rpath,$ORIGIN/../../math_amd"]
        if verbose:
            print(" ".join(cmd))
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"C++ test build failed\n{r.stdout}\n{r.stderr}")
        exes.append(exe)
    return exes[0]


if __name__ == "__main__":
    v = "-v" in sys.argv
    print(build_library(v))
    print(build_oracle(v))
    print(build_cpp_tests(v))
