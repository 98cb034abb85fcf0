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

HIP_SOURCES = ["atm_kernels.hip", "atm_backward.hip", "atm_restraints.hip", "atm_md.hip", "atm_pme.hip",
               "atm_noalloc.cpp",
               "atm_capi.cpp", "atm_pme_capi.cpp"]


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


CXX = os.environ.get("CXX", "g++")
CXX_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-parameter", "-fvisibility=hidden"]
CORE_INC = ["-I" + os.path.join(PKG, "openmm_min", "include")]
ATM_INC = CORE_INC + ["-I" + os.path.join(PKG, d) for d in
                      ("openmmapi/include", "serialization/include", "platforms/mi355x/include")] + ["-I" + INCLUDE]


def _sources(*dirs):
    out = []
    for d in dirs:
        full = os.path.join(PKG, d)
        out += sorted(os.path.join(full, f) for f in os.listdir(full) if f.endswith(".cpp"))
    return out


def build_cpp_layer(jobs: int = 4) -> None:
    """libOpenMMMin.so (minimal OpenMM-compatible core) and libATMMetaForcePlugin.so
    (ATMMetaForce API, XML proxy, MI355X platform) -- host C++ built with g++."""
    os.makedirs(BUILD, exist_ok=True)

    def compile_one(args):
        src, inc = args
        obj = os.path.join(BUILD, os.path.basename(os.path.dirname(os.path.dirname(src))) + "_" +
                           os.path.splitext(os.path.basename(src))[0] + ".o")
        _run([CXX] + CXX_FLAGS + inc + ["-c", src, "-o", obj])
        return obj

    core_src = _sources("openmm_min/src")
    atm_src = _sources("openmmapi/src", "serialization/src", "platforms/mi355x/src")
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        core_obj = list(ex.map(compile_one, [(s, CORE_INC) for s in core_src]))
        atm_obj = list(ex.map(compile_one, [(s, ATM_INC) for s in atm_src]))
    core = os.path.join(LIB, "libOpenMMMin.so")
    _run([CXX, "-shared", "-o", core + ".tmp"] + core_obj + ["-ldl"])
    os.replace(core + ".tmp", core)
    plugin = os.path.join(LIB, "libATMMetaForcePlugin.so")
    _run([CXX, "-shared", "-o", plugin + ".tmp"] + atm_obj +
         ["-L" + LIB, "-lOpenMMMin", "-latm_mi355x", "-Wl,-rpath,$ORIGIN"])
    os.replace(plugin + ".tmp", plugin)


def build_python_module() -> str:
    """pybind11 module atmmetaforce/_atmmetaforce*.so (links the two libraries above)."""
    import pybind11
    suffix = sysconfig.get_config_var("EXT_SUFFIX")
    out = os.path.join(PKG, "atmmetaforce", "_atmmetaforce" + suffix)
    inc = ATM_INC + ["-I" + pybind11.get_include(), "-I" + sysconfig.get_paths()["include"]]
    src = os.path.join(PKG, "python", "atmmetaforce_module.cpp")
    _run([CXX] + CXX_FLAGS + inc + ["-shared", src, "-o", out + ".tmp", "-L" + LIB, "-lATMMetaForcePlugin",
                                    "-lOpenMMMin", "-latm_mi355x", "-Wl,-rpath,$ORIGIN/../lib"])
    os.replace(out + ".tmp", out)
    return out


def build_all(jobs: int = 4) -> None:
    build_hip_library(jobs)
    build_cpp_layer(jobs)
    build_python_module()


if __name__ == "__main__":
    build_all()
    print("built", os.path.join(LIB, "libatm_mi355x.so"))
