"""Build the HIP library ``libbeta1_mhd.so`` in-tree for gfx950.

The host set-up sources are compiled with g++ (OpenMP), the kernels and the
C-ABI with hipcc.  Every object is rebuilt on each call so that a build is
always a real compile check.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_objs")
LIB = os.path.join(HERE, "libbeta1_mhd.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-ffp-contract=off", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function"]
HIP_FLAGS = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-munsafe-fp-atomics"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"compile failed: {cmd[-1]}")
    return r


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    jobs = []
    for src in ("host/setup.cpp", "host/cases.cpp", "host/host_capi.cpp", "host/formats.cpp",
                "host/tecplot.cpp", "host/dropin.cpp"):
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        jobs.append(["g++", *COMMON, "-fopenmp", "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in ("device/kernels.hip", "device/capi.cpp"):
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        jobs.append([HIPCC, *COMMON, *HIP_FLAGS, *lang, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    procs = [subprocess.Popen(j, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True) for j in jobs]
    failed = []
    for j, p in zip(jobs, procs):
        out, err = p.communicate()
        if p.returncode != 0:
            failed.append((j, out + err))
        elif verbose and (out or err):
            sys.stderr.write(out + err)
    if failed:
        for j, msg in failed:
            sys.stderr.write(" ".join(j) + "\n" + msg + "\n")
        raise RuntimeError("libbeta1_mhd build failed")
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fopenmp", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
