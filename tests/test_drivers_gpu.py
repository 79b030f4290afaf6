"""The reference-compatible drivers on the GPU: every pair and V verifies
every output element against the driver's host loop and reports the
reference's result lines."""
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu

PKG = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "magicube_amd")


@pytest.mark.parametrize("preA,preB,V", [(8, 8, 8), (8, 8, 4), (8, 8, 2), (16, 8, 8), (16, 8, 2), (8, 4, 8),
                                         (8, 4, 4), (4, 4, 8), (4, 4, 2), (12, 8, 8), (12, 8, 2), (16, 16, 8),
                                         (16, 16, 4), (12, 4, 8), (12, 4, 2), (16, 4, 8), (16, 4, 4)])
def test_spmm_driver_passes(preA, preB, V):
    cmd = [os.path.join(PKG, "spmm_benchmark"), "synthetic:256,512,0.8,11", "128", str(V), "0", "1", "1", "1",
           str(preA), str(preB)]
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Results verification: PASS" in r.stdout, r.stdout
    assert re.search(r"Magicube SpMM runtime [0-9.e-]+ ms", r.stdout)
    assert re.search(r"performance TOP/s: [0-9.e+-]+", r.stdout)


@pytest.mark.parametrize("bits,K,V", [(8, 64, 8), (8, 128, 4), (4, 64, 8), (4, 256, 2), (16, 64, 8), (16, 256, 2)])
def test_sddmm_driver_passes(bits, K, V):
    cmd = [os.path.join(PKG, "sddmm_benchmark"), "synthetic:256,384,0.8,13", str(K), str(V), "1", "1", "1",
           str(bits), str(bits)]
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "SDDMM results verification: PASS" in r.stdout, r.stdout
    assert re.search(r"Magicube SDDMM runtime [0-9.e-]+ ms", r.stdout)


# the reference's launch_spmm_*.py / launch_sddmm_*.py modes (sparse, kernel, mixed)
@pytest.mark.parametrize("sparse,kernel,V", [(-1, 0, 8), (-2, 0, 4), (3, 0, 8), (3, 0, 2), (2, 0, 8), (1, 3, 4)])
def test_baseline_spmm_driver_verifies(sparse, kernel, V):
    cmd = [os.path.join(PKG, "baseline_spmm_benchmark"), "synthetic:256,512,0.8,11", "128", str(V), str(kernel),
           "1", "1", str(sparse), "1"]
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Results verified: they agree." in r.stdout, r.stdout
    assert re.search(r"runtime [0-9.e-]+ ms", r.stdout)


@pytest.mark.parametrize("sparse,kernel,V", [(0, 0, 8), (-1, 0, 4), (1, 3, 8)])
def test_baseline_sddmm_driver_verifies(sparse, kernel, V):
    cmd = [os.path.join(PKG, "baseline_sddmm_benchmark"), "synthetic:256,384,0.8,13", "64", str(V), str(kernel),
           "2", "1", "1", str(sparse), "1"]
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Results verified: they agree." in r.stdout, r.stdout
    assert re.search(r"runtime [0-9.e-]+ ms", r.stdout)
