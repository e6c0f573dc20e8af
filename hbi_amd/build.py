"""Build libhbi.so (HIP, gfx950) and the Fortran host driver in-tree.

    python -m hbi_amd.build          # library + Fortran driver
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
FSRC = os.path.join(HERE, "fortran")
LIB = os.path.join(HERE, "libhbi.so")
DRIVER = os.path.join(HERE, "lhbiem_gpu")
DRIVER_FV = os.path.join(HERE, "lhbiem_fv_gpu")
DRIVER_FV3 = os.path.join(HERE, "lhbiem_fv3dp_gpu")

SOURCES = ["partition.cpp", "hmat.hip", "rk.hip", "multirhs.hip", "ens.hip", "learn.hip", "fv.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FC = os.environ.get("FC_HBI", "/opt/rocm/bin/amdflang")
ARCH = os.environ.get("HBI_ARCH", "gfx950")

# Packed FP32 instructions are disabled for all device code.
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
             "-Wno-unused-variable", "-Wno-unused-but-set-variable", "-munsafe-fp-atomics",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-I" + os.path.join(ROOT, "include")]


def _run(cmd, cwd=None):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def build_lib(jobs: int = 4) -> str:
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    procs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(objdir, src + ".o")
        objs.append(obj)
        cmd = [HIPCC] + HIP_FLAGS + ["-c", path, "-o", obj]
        if src.endswith(".cpp"):
            cmd = [HIPCC, "-O2", "-fPIC", "-std=c++17", "-Wall", "-I" + os.path.join(ROOT, "include"), "-c", path,
                   "-o", obj]
        print(" ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
        if len(procs) >= jobs:
            if procs.pop(0).wait() != 0:
                raise RuntimeError("compile failed")
    for p in procs:
        if p.wait() != 0:
            raise RuntimeError("compile failed")
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


def build_driver() -> str:
    """Fortran hosts (amdflang) linked against libhbi.so: lhbiem_gpu (main_LH),
    lhbiem_fv_gpu (main_fv) and lhbiem_fv3dp_gpu (main_fv3dp)."""
    srcs = ["m_hbi_abi.f90", "m_hacapk_gpu.f90", "lhbiem_gpu.f90", "lhbiem_fv_gpu.f90", "lhbiem_fv3dp_gpu.f90"]
    paths = [os.path.join(FSRC, s) for s in srcs]
    if not all(os.path.exists(p) for p in paths) or not shutil.which(FC) and not os.path.exists(FC):
        return ""
    objdir = os.path.join(HERE, "_build", "f")
    os.makedirs(objdir, exist_ok=True)
    objs = {}
    for p in paths:
        o = os.path.join(objdir, os.path.basename(p) + ".o")
        _run([FC, "-O2", "-c", p, "-o", o, "-module-dir", objdir, "-I", objdir])
        objs[os.path.basename(p)] = o
    common = [objs["m_hbi_abi.f90"], objs["m_hacapk_gpu.f90"]]
    link = ["-L" + HERE, "-lhbi", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + HERE]
    _run([FC, "-o", DRIVER] + common + [objs["lhbiem_gpu.f90"]] + link)
    _run([FC, "-o", DRIVER_FV] + common + [objs["lhbiem_fv_gpu.f90"]] + link)
    _run([FC, "-o", DRIVER_FV3] + common + [objs["lhbiem_fv3dp_gpu.f90"]] + link)
    return DRIVER


def main(argv=None):
    build_lib()
    build_driver()


if __name__ == "__main__":
    main(sys.argv[1:])
