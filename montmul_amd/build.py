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


def _run_all(cmds, verbose):
    """Runs the commands in parallel; fails with the first failing one's output."""
    procs = []
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    failed = None
    for cmd, p in procs:
        text = p.communicate()[0]
        out = "\n".join(l for l in text.splitlines() if "packed-fp32-ops" not in l)
        if p.returncode != 0 and failed is None:
            failed = f"build step failed ({p.returncode}): {' '.join(cmd)}\n{out}"
        elif verbose and out.strip():
            print(out)
    if failed:
        raise RuntimeError(failed)


def build(verbose=False, out=LIB, extra=()):
    os.makedirs(OUT_DIR, exist_ok=True)
    hip_obj = os.path.join(OUT_DIR, "mm_hip.o")
    probe_obj = os.path.join(OUT_DIR, "mm_probe.o")
    host_obj = os.path.join(OUT_DIR, "mm_host.o")
    serve_obj = os.path.join(OUT_DIR, "mm_serve.o")
    multi_obj = os.path.join(OUT_DIR, "mm_multi.o")
    # The mixed-length kernel is built once per workgroup size, each in an object of its own:
    # the three device compiles run side by side.
    mix_objs = [os.path.join(OUT_DIR, f"mm_mix{nt}.o") for nt in (512, 1024)]
    _run_all([[HIPCC, *HIP_FLAGS, *extra, "-c", os.path.join(CSRC, "mm_hip.hip"), "-o", hip_obj],
              [HIPCC, *HIP_FLAGS, *extra, "-c", os.path.join(CSRC, "mm_probe.hip"), "-o", probe_obj]] +
             [[HIPCC, *HIP_FLAGS, *extra, f"-DMM_MIX_NT={nt}", "-c", os.path.join(CSRC, "mm_mix.hip"), "-o", o]
              for nt, o in zip((512, 1024), mix_objs)] +
             [["g++", "-O2", "-std=c++17", "-fPIC", "-Wall", "-pthread", f"-I{GMP_INCLUDE}", "-c",
               os.path.join(CSRC, "mm_host.cpp"), "-o", host_obj],
              ["g++", "-O2", "-std=c++17", "-fPIC", "-Wall", "-pthread", f"-I{GMP_INCLUDE}", "-c",
               os.path.join(CSRC, "mm_serve.cpp"), "-o", serve_obj],
              ["g++", "-O2", "-std=c++17", "-fPIC", "-Wall", "-pthread", f"-I{GMP_INCLUDE}", "-c",
               os.path.join(CSRC, "mm_multi.cpp"), "-o", multi_obj]], verbose)
    tmp = out + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp, hip_obj, probe_obj, *mix_objs, host_obj, serve_obj, multi_obj,
          GMP_LIB, "-pthread"], verbose)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
