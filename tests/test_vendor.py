"""Vendor comparator library on the CPU: it loads and exports every symbol its
header declares, rejects bad arguments before any device call, and the
format converters describe the same matrix as the oracle's vector-sparse
product."""
import ctypes
import os
import re

import numpy as np
import pytest

from magicube_amd import synthetic, vendor

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "magicube_vendor.h")).read()
    return sorted(set(re.findall(r"\b(mcv_\w+)\s*\(", text)))


def test_exports_every_declared_symbol():
    lib = vendor.lib()
    names = declared()
    assert names and set(names) == set(vendor.exported_symbols())
    for name in names:
        assert hasattr(lib, name), name


def test_argument_errors_before_device_calls():
    lib = vendor.lib()
    h = ctypes.c_void_p()
    p = ctypes.c_void_p(16)
    # shape, dtype, algorithm and pointer checks return before any HIP call
    assert lib.mcv_csr_spmm_create(0, 0, 0, 8, 8, 4, p, p, p, p, p, None, ctypes.byref(h)) == 2
    assert lib.mcv_csr_spmm_create(7, 0, 8, 8, 8, 4, p, p, p, p, p, None, ctypes.byref(h)) == 3
    assert lib.mcv_csr_spmm_create(0, 9, 8, 8, 8, 4, p, p, p, p, p, None, ctypes.byref(h)) == 3
    assert lib.mcv_csr_spmm_create(0, 0, 8, 8, 8, 4, None, p, p, p, p, None, ctypes.byref(h)) == 3
    assert lib.mcv_csr_spmm_create(0, 0, 8, 8, 8, 4, p, p, p, p, p, None, None) == 3
    assert lib.mcv_bsr_spmm_create(0, 2, 2, 8, 4, 0, p, p, p, p, p, None, ctypes.byref(h)) == 2
    assert lib.mcv_csr_sddmm_create(8, 8, 0, 4, p, p, p, p, p, None, ctypes.byref(h)) == 2
    assert lib.mcv_gemm_create(0, 8, 8, 8, p, None, 0, p, None, ctypes.byref(h)) == 3
    assert lib.mcv_gemm_create(2, 8, 8, 8, p, p, 0, p, None, ctypes.byref(h)) == 3
    assert lib.mcv_run(None) == 3
    assert b"mcv_run" in lib.mcv_last_error()
    assert lib.mcv_workspace_bytes(None) == 0
    assert lib.mcv_destroy(None) == 0


@pytest.mark.parametrize("V", [2, 4, 8])
def test_converters_describe_the_oracle_matrix(orc, V):
    csr = synthetic.poisson_block_mask(48 * V, 160, V, 0.8, seed=V)
    a = synthetic.random_lanes(csr.nnz_vec * V, 8, True, 1).astype(np.int8)
    b = synthetic.random_lanes(160 * 24, 8, True, 2)
    dense = vendor.dense_lhs(csr, a)
    exp = orc.spmm_plain(csr.m_vec, V, 24, csr.row_offsets, csr.col_indices, a, b)
    assert np.array_equal(dense.astype(np.int64) @ b.reshape(160, 24).astype(np.int64), exp)

    rp, ci, gather = vendor.scalar_csr(csr)
    assert rp[0] == 0 and rp[-1] == csr.nnz_vec * V and np.all(np.diff(rp) >= 0)
    d2 = np.zeros_like(dense)
    rows = np.repeat(np.arange(csr.m), np.diff(rp))
    d2[rows, ci] = a[gather]
    assert np.array_equal(d2, dense)
    for r in range(csr.m):   # ascending columns within each scalar row
        assert np.all(np.diff(ci[rp[r]:rp[r + 1]]) > 0)

    brp, bci, bv = vendor.bsr_blocks(csr, a)
    blocks = bv.reshape(-1, V, V)
    d3 = np.zeros_like(dense)
    for i in range(csr.m_vec):
        assert np.all(np.diff(bci[brp[i]:brp[i + 1]]) > 0)
        for t in range(brp[i], brp[i + 1]):
            d3[i * V:(i + 1) * V, bci[t] * V:(bci[t] + 1) * V] = blocks[t]
    assert np.array_equal(d3, dense)

    idx = vendor.vector_to_scalar_index(csr)
    assert np.array_equal(a[gather][idx], a)


def test_bsr_needs_k_multiple_of_v():
    csr = synthetic.poisson_block_mask(32, 30, 4, 0.5, seed=1)
    with pytest.raises(ValueError):
        vendor.bsr_blocks(csr, np.zeros(csr.nnz_vec * 4, np.int8))


@pytest.mark.parametrize("density", [0.02, 0.1, 0.3])
def test_uniform_blocks(density):
    m_vec, k, V = 40, 2048, 8
    rp, ci = vendor.uniform_blocks(m_vec, k, V, density, seed=5)
    per_row = int(density * k / V) + 1
    assert np.array_equal(rp, np.arange(m_vec + 1) * per_row)
    cols = ci.reshape(m_vec, per_row)
    assert np.all(np.diff(cols, axis=1) > 0)
    assert cols.min() >= 0 and cols.max() < k // V
    # rows differ: the columns are drawn per block row
    assert len({tuple(r) for r in cols}) > m_vec // 2
