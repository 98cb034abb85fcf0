"""Build the native parts of the package in-tree.

  lib/libatm_mi355x.so     gfx950 kernels + the C ABI of include/atm_mi355x.h (hipcc)
  lib/libATMMetaForcePlugin.so
                           the ATMMetaForce C++ API, its ForceImpl, the XML proxy,
                           the minimal OpenMM-compatible core and the MI355X
                           platform kernels (g++, links libatm_mi355x.so)
  atmmetaforce/_atmmetaforce*.so
                           pybind11 module mirroring the reference's SWIG module

Every HIP translation unit is compiled for gfx950 only, with packed-FP32
instruction selection disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
import sysconfig

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "lib")
BUILD = os.path.join(PKG, "_build")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + CSRC, "-I" + INCLUDE]

HIP_SOURCES = ["atm_kernels.hip", "atm_backward.hip", "atm_capi.cpp"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + out)
    return out


def build_hip_library(jobs: int = 4) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)

    def compile_one(src):
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".cpp") else []
        _run([HIPCC] + HIP_FLAGS + lang + ["-c", os.path.join(CSRC, src), "-o", obj])
        return obj

    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, HIP_SOURCES))
    out = os.path.join(LIB, "libatm_mi355x.so")
    tmp = out + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, out)
    return out


def build_all(jobs: int = 4) -> None:
    build_hip_library(jobs)
    try:
        import plugin_build  # noqa: F401  (C++ API layer, optional until present)
    except ImportError:
        pass


if __name__ == "__main__":
    build_all()
    print("built", os.path.join(LIB, "libatm_mi355x.so"))
