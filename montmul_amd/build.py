"""Builds montmul_amd/libmontmul_hip.so in the source tree (hipcc for gfx950, g++ for the host).

Usage: python -m montmul_amd.build   (or montmul_amd.build.build() from Python)
"""

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmontmul_hip.so")
# GMP headers come from the conda prefix; the library is the system's libgmp.so.10 (the same
# 6.2.1 release), linked by path so that no rpath drags conda's older libstdc++ into processes
# that also load the HIP runtime.
GMP_INCLUDE = os.environ.get("MONTMUL_GMP_INCLUDE", "/opt/conda/include")
GMP_LIB = os.environ.get("MONTMUL_GMP_LIB", "/usr/lib/x86_64-linux-gnu/libgmp.so.10")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# Packed FP32 instructions are kept out of every device build.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result", "-Wno-unused-value"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    if verbose and out.strip():
        print(out)


def build(verbose=False, out=LIB, extra=()):
    os.makedirs(OUT_DIR, exist_ok=True)
    hip_obj = os.path.join(OUT_DIR, "mm_hip.o")
    host_obj = os.path.join(OUT_DIR, "mm_host.o")
    _run([HIPCC, *HIP_FLAGS, *extra, "-c", os.path.join(CSRC, "mm_hip.hip"), "-o", hip_obj], verbose)
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-Wall", f"-I{GMP_INCLUDE}", "-c",
          os.path.join(CSRC, "mm_host.cpp"), "-o", host_obj], verbose)
    serve_obj = os.path.join(OUT_DIR, "mm_serve.o")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-Wall", "-pthread", f"-I{GMP_INCLUDE}", "-c",
          os.path.join(CSRC, "mm_serve.cpp"), "-o", serve_obj], verbose)
    tmp = out + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp, hip_obj, host_obj, serve_obj,
          GMP_LIB, "-pthread"], verbose)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
