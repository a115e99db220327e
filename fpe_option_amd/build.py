"""Build libfpe_engine.so (HIP kernels for gfx950 + host engine + C ABI) in-tree.

Usage: python -m fpe_option_amd.build  (or fpe_option_amd.build.build()).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libfpe_engine.so")
LU_BENCH = os.path.join(OUT_DIR, "lu_bench")
ARCH = os.environ.get("FPE_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

# Packed FP32 VALU forms are disabled for every device compile.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-Wno-unused-result", "-I", os.path.join(HERE, "..", "include")]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + out)
    return out


def build(verbose=False):
    os.makedirs(OUT_DIR, exist_ok=True)
    objs = []
    for src in ("domain.cpp", "capi.cpp"):
        o = os.path.join(OUT_DIR, src + ".o")
        _run([CXX] + CXX_FLAGS + ["-c", os.path.join(CSRC, src), "-o", o])
        objs.append(o)
    for src in ("engine.hip", "surrogate.hip", "pod.hip"):
        o = os.path.join(OUT_DIR, src + ".o")
        msg = _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", o])
        if verbose and msg:
            print(msg)
        objs.append(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    # the factorisation microbenchmark (a measurement tool for profiles/, not part of the library)
    tmp = LU_BENCH + ".tmp"
    _run([HIPCC] + HIP_FLAGS + [os.path.join(CSRC, "lu_bench.hip"), "-o", tmp])
    os.replace(tmp, LU_BENCH)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
