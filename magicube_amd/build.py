"""Build recipe for the gfx950 library: ``python -m magicube_amd.build``.

Compiles every translation unit under ``magicube_amd/csrc`` with hipcc for
gfx950 (in parallel, into ``magicube_amd/_build``) and links them into the
in-tree shared library ``magicube_amd/libmagicube_hip.so``; links the
vendor-library comparators into ``magicube_amd/libmagicube_vendor.so`` and
builds the reference-compatible benchmark drivers when their sources are
present.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmagicube_hip.so")
ARCH = os.environ.get("MC_OFFLOAD_ARCH", "gfx950")

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = shutil.which("hipcc") or f"{ROCM}/bin/hipcc"
# Packed FP32 instructions stay disabled for every device pass.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Wall", "-Wno-unused-variable"]

LIB_SOURCES = [
    "common.cpp",
    "format.cpp",
    "spmm_api.hip",
    "spmm_l8r8.hip",
    "spmm_l16r8.hip",
    "spmm_l8r4.hip",
    "spmm_l4r4.hip",
    "spmm_l12r8.hip",
    "spmm_l16r16.hip",
    "spmm_l12r4.hip",
    "spmm_l16r4.hip",
]
OPTIONAL_SOURCES = ["sddmm.hip", "attention.hip", "layernorm.hip"]
# Vendor-library comparators (rocSPARSE, hipBLASLt): a library of their own so
# the product library carries no dependency on them.
VENDOR_LIB = os.path.join(HERE, "libmagicube_vendor.so")
VENDOR_SOURCES = ["vendor.cpp"]
VENDOR_LIBS = ["-lrocsparse", "-lhipblaslt", "-lamdhip64"]
DRIVERS = {"spmm_benchmark": "spmm_benchmark.cpp", "sddmm_benchmark": "sddmm_benchmark.cpp"}
# the reference's vendor-library baseline drivers, on rocSPARSE / hipBLASLt
VENDOR_DRIVERS = {"baseline_spmm_benchmark": "baseline_spmm_benchmark.cpp",
                  "baseline_sddmm_benchmark": "baseline_sddmm_benchmark.cpp"}
# measurement programs with device code of their own
HIP_TOOLS = {"gather_floor": "gather_floor.hip"}


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    path = os.path.join(CSRC, src)
    if src.endswith(".hip"):
        cmd = [HIPCC, *COMMON, *DEVICE_FLAGS, "-x", "hip", "-c", path, "-o", obj]
    else:
        cmd = [HIPCC, "-O3", "-std=c++17", "-fPIC", "-Wall", "-x", "c++", "-D__HIP_PLATFORM_AMD__",
               f"-I{ROCM}/include", "-c", path, "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj


def build(jobs: int | None = None, verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    sources = LIB_SOURCES + [s for s in OPTIONAL_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, sources))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    vobjs = [_compile(src) for src in VENDOR_SOURCES]
    tmp = VENDOR_LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "-o", tmp, *vobjs, f"-L{ROCM}/lib", f"-Wl,-rpath,{ROCM}/lib", *VENDOR_LIBS]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, VENDOR_LIB)
    drivers = [(exe, src, []) for exe, src in DRIVERS.items()]
    drivers += [(exe, src, ["-lmagicube_vendor"]) for exe, src in VENDOR_DRIVERS.items()]
    for exe, src, extra in drivers:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        out = os.path.join(HERE, exe)
        cmd = [HIPCC, "-O3", "-std=c++17", "-x", "c++", "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include",
               path, "-o", out, f"-L{HERE}", *extra, "-lmagicube_hip", "-Wl,-rpath,$ORIGIN", f"-L{ROCM}/lib",
               f"-Wl,-rpath,{ROCM}/lib", "-lamdhip64"]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"driver build failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    for exe, src in HIP_TOOLS.items():
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        cmd = [HIPCC, *COMMON, *DEVICE_FLAGS, "-x", "hip", path, "-o", os.path.join(HERE, exe)]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"tool build failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
