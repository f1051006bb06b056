"""Builds the native libraries in-tree.

libbfgpu.so  — HIP kernels + C ABI (hipcc, gfx950)        -> butterfly-osm_amd/lib/
               (bfgpu.hip: PHAST engine and C ABI; meet.hip: the many-to-many join;
                customize.hip: triangle relaxation of the CCH)
libbfcch.so  — synthetic road network -> CCH builder (g++)  -> butterfly-osm_amd/lib/
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "lib")
BUILD = os.path.join(PKG, "_build")
REPO = os.path.dirname(PKG)

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# packed FP32 instructions are kept off for every HIP translation unit
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument"]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-sign-compare"]


def _run(cmd):
    print("[build]", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_bfgpu():
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(BUILD, exist_ok=True)
    lay_o = os.path.join(BUILD, "layout.o")
    _run(["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, "layout.cpp"), "-o", lay_o])
    objs = []
    for name in ("bfgpu", "meet", "customize"):
        o = os.path.join(BUILD, name + ".o")
        _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, name + ".hip"), "-o", o])
        objs.append(o)
    out = os.path.join(LIB, "libbfgpu.so")
    tmp = out + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, lay_o, "-o", tmp])
    os.replace(tmp, out)
    return out


def build_cchbuild():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libbfcch.so")
    tmp = out + ".tmp"
    _run(["g++", *CXX_FLAGS, "-fopenmp", "-shared",
          os.path.join(CSRC, "cchbuild", "cchbuild.cpp"), "-o", tmp])
    os.replace(tmp, out)
    return out


def build_all():
    return [build_cchbuild(), build_bfgpu()]


if __name__ == "__main__":
    build_all()
    sys.exit(0)
