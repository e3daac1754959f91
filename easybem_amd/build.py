"""In-tree build of the HIP library (libeasybem_hip.so) and the pybind11
drop-in module ``easybem`` for gfx950.

    python -m easybem_amd.build

Every HIP translation unit is compiled with packed FP32 instructions
disabled (-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
import sysconfig

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libeasybem_hip.so")
ARCH = os.environ.get("EBH_OFFLOAD_ARCH", "gfx950")

HIPCC = "/opt/rocm/bin/hipcc"
COMMON = ["-O3", "-std=c++17", "-fPIC", "-I" + os.path.join(ROOT, "include"),
          "-I" + CSRC, "-Wall", "-Wno-unused-function"]
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops", "-fgpu-flush-denormals-to-zero"]

LIB_SOURCES = ["host_geometry.cpp", "near_field.hip", "far_field.hip",
               "capi.hip", "solver.hip", "legacy_api.cpp", "legacy_dense.cpp", "server.cpp",
               "autotune.cpp", "hypersingular.hip", "evaluate.hip",
               "dense.hip", "elasticity.hip", "spai.hip"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, src + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC, "-x", "hip", *COMMON, *DEVICE, "-c", path, "-o", obj]
    else:
        cmd = [HIPCC, *COMMON, "-c", path, "-o", obj]
    _run(cmd)
    return obj


def build_lib(jobs: int = 8) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in LIB_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, *objs, "-o", tmp,
          "-lpthread", "-lrocblas"])
    os.replace(tmp, LIB)
    return LIB


def build_module() -> str:
    """pybind11 module ``easybem`` (drop-in for the reference module)."""
    import pybind11
    src = os.path.join(CSRC, "bindings.cpp")
    if not os.path.exists(src):
        return ""
    ext = sysconfig.get_config_var("EXT_SUFFIX")
    out = os.path.join(PKG, "easybem" + ext)
    py_inc = sysconfig.get_paths()["include"]
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared",
          "-I" + pybind11.get_include(), "-I" + py_inc,
          "-I" + os.path.join(ROOT, "include"), src, "-o", tmp,
          "-L" + PKG, "-leasybem_hip", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, out)
    return out


def build_cpp_check() -> str:
    """tests/cpp/dropin_check: a C++ caller of the drop-in dense surface
    (include/BEM3dLaplace.h, include/BEM3dElasticity.h) linked against the
    library, run by the tests."""
    src = os.path.join(ROOT, "tests", "cpp", "dropin_check.cpp")
    if not os.path.exists(src):
        return ""
    out = os.path.join(ROOT, "tests", "cpp", "dropin_check")
    _run(["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"), src,
          "-o", out + ".tmp", "-L" + PKG, "-leasybem_hip",
          "-Wl,-rpath,$ORIGIN/../../easybem_amd"])
    os.replace(out + ".tmp", out)
    return out


def build(jobs: int = 8) -> None:
    build_lib(jobs)
    build_module()
    build_cpp_check()


if __name__ == "__main__":
    build(int(sys.argv[1]) if len(sys.argv) > 1 else 8)
    print(LIB)
