"""Reference-compatible drivers and the C++ compat shim, on the CPU: the
shim compiles against the headers and maps rejected arguments to
hipErrorInvalidValue; the drivers parse .smtx and synthetic patterns and
print the reference's header lines with the alignment the format code
computes (the device part needs a GPU and is covered by the gpu tests)."""
import os
import re
import subprocess

import numpy as np
import pytest

from magicube_amd import formats, synthetic

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "magicube_amd")
SPMM = os.path.join(PKG, "spmm_benchmark")
SDDMM = os.path.join(PKG, "sddmm_benchmark")


def _need(path):
    if not os.path.exists(path):
        pytest.skip(f"{path} not built (python -m magicube_amd.build)")


def write_smtx(path, csr):
    with open(path, "w") as f:
        f.write(f"{csr.m_vec}, {csr.k}, {csr.nnz_vec}\n")
        f.write(" ".join(map(str, csr.row_offsets.tolist())) + "\n")
        f.write(" ".join(map(str, csr.col_indices.tolist())) + "\n")


def test_compat_shim_compiles_and_maps_errors(tmp_path):
    src = tmp_path / "shim.cpp"
    src.write_text(r'''
#include "magicube_compat.hpp"
#include <cstdio>
int main() {
    int dummy[4] = {0, 0, 0, 0};
    hipError_t a = spmm::wmmaSpmm_8b(4, 3, 16, 16, nullptr, dummy, dummy, dummy, dummy, dummy);
    hipError_t b = sddmm::wmmaSddmm_8b(4, 63, 16, nullptr, dummy, dummy, dummy, dummy, dummy, 8);
    hipError_t c = spmm::wmmaSpmm_4b(0, 8, 16, 16, nullptr, dummy, dummy, dummy, dummy, dummy);
    std::printf("%d %d %d\n", (int)a, (int)b, (int)c);
    return 0;
}
''')
    exe = tmp_path / "shim"
    cmd = ["g++", "-std=c++17", "-D__HIP_PLATFORM_AMD__", f"-I{ROOT}/include", "-I/opt/rocm/include", str(src),
           "-o", str(exe), f"-L{PKG}", "-lmagicube_hip", f"-Wl,-rpath,{PKG}", "-L/opt/rocm/lib", "-lamdhip64"]
    _need(os.path.join(PKG, "libmagicube_hip.so"))
    subprocess.run(cmd, check=True, capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    invalid = 1   # hipErrorInvalidValue
    assert out == [str(invalid), str(invalid), "0"], out   # m_vec = 0 is a no-op


def _header(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=60)
    return r.stdout


def test_spmm_driver_reads_smtx_and_prints_reference_lines(tmp_path):
    _need(SPMM)
    csr = synthetic.poisson_block_mask(256, 384, 8, 0.8, seed=3)
    path = tmp_path / "m.smtx"
    write_smtx(path, csr)
    for preA, preB, align in ((8, 8, 16), (4, 4, 32), (12, 8, 16), (16, 16, 16), (12, 4, 32)):
        out = _header([SPMM, str(path), "64", "8", "0", "1", "1", "1", str(preA), str(preB)])
        assert out.startswith(f"Sparse matrix: {path}\n")
        # the reference prints the container width, then the value width
        assert f"preA {16 if preA == 12 else preA}, preA_cut {preA}, preB {preB}, vec_length 8 \n" in out
        assert f"m_vec 32, dimN 64, nonzeros_vec {csr.nnz_vec}, dimk 384, mma_k_dim {align} \n" in out
        _, an = formats.align_rows(csr.row_offsets, align)
        assert f" nonzero_vec: {csr.nnz_vec} aligned_ nonzero_vec: {an}\n" in out
        assert f"total Gflops: " in out


def test_drivers_reject_unsupported_combinations():
    _need(SPMM)
    _need(SDDMM)
    assert "Unsupported precision and vec_length!" in _header([SPMM, "x", "64", "8", "0", "1", "1", "1", "8", "16"])
    assert "Unsupported precision and vec_length!" in _header([SPMM, "x", "64", "8", "0", "1", "1", "1", "12", "16"])
    assert "Unsupported Kernel" in _header([SPMM, "x", "64", "8", "2", "1", "1", "1", "8", "8"])
    assert "Unsupported precision!" in _header([SDDMM, "x", "64", "8", "1", "1", "1", "16", "8"])
    assert "Unsupported vec_length!" in _header([SDDMM, "x", "64", "3", "1", "1", "1", "8", "8"])


def test_sddmm_driver_synthetic_header():
    _need(SDDMM)
    out = _header([SDDMM, "synthetic:512,256,0.9,5", "64", "8", "1", "1", "1", "8", "8"])
    m = re.search(r"PreA: 8, PreB: 8, vec_len: 8, M: 512, M_vec: 64, N: 256, nnz: (\d+), K: 64", out)
    assert m and int(m.group(1)) % 8 == 0
    m2 = re.search(r" nonzero_vec: (\d+) aligned_ nonzero_vec: (\d+)", out)
    assert m2 and int(m2.group(1)) * 8 == int(m.group(1)) and int(m2.group(2)) % 8 == 0


def test_baseline_drivers_usage_and_unsupported_modes():
    """The vendor-baseline drivers print usage, and reject the reference's own
    vectorSparse modes before any device call."""
    for exe, argv in (("baseline_spmm_benchmark", ["synthetic:64,128,0.8,3", "64", "8", "0", "1", "0", "1", "1"]),
                      ("baseline_sddmm_benchmark", ["synthetic:64,128,0.8,3", "64", "8", "0", "2", "1", "0", "1",
                                                    "1"])):
        path = os.path.join(PKG, exe)
        _need(path)
        r = subprocess.run([path, "--help"], capture_output=True, text=True, timeout=60)
        assert r.returncode == 0 and "usage" in r.stdout
        r = subprocess.run([path, *argv], capture_output=True, text=True, timeout=60)
        assert r.returncode == 2 and "not carried" in r.stderr, r.stdout + r.stderr
        assert "Sparse matrix: synthetic:64,128,0.8,3" in r.stdout
