"""Vendor-library comparators on the benchmark operands: rocSPARSE SpMM (CSR,
BSR) and SDDMM, and hipBLASLt dense GEMM, through ``libmagicube_vendor.so``
(include/magicube_vendor.h).

They stand where the reference's baselines call cuSPARSE and cuBLAS:
Blocked-ELL SpMM with V x V blocks at the vector matrix's density, int8 and
fp16 (baselines/spmm_benchmark.cpp:268-602; here BSR, as rocSPARSE's
Blocked-ELL SpMM has neither type), dense int8 and fp16 GEMM
(baselines/spmm_benchmark.cpp:603-770, baselines/sddmm_benchmark.cpp:362-430)
and SDDMM (baselines/sddmm_benchmark.cpp:33-88).  The converters below also
give the vendor formats of the very same vector-sparse matrix (scalar CSR, BSR
with each vector padded to a V x V block).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .formats import VectorCSR

HERE = os.path.dirname(os.path.abspath(__file__))
VENDOR_PATH = os.path.join(HERE, "libmagicube_vendor.so")

I8 = 0
F16 = 1
CSR_ALGS = {"default": 0, "row_split": 1, "nnz_split": 2, "merge_path": 3}

_P = ctypes.c_void_p
_I = ctypes.c_int
_LL = ctypes.c_longlong
_SIGS = {
    "mcv_csr_spmm_create": (_I, [_I, _I, _LL, _LL, _LL, _LL, _P, _P, _P, _P, _P, _P, _P]),
    "mcv_bsr_spmm_create": (_I, [_I, _LL, _LL, _LL, _LL, _I, _P, _P, _P, _P, _P, _P, _P]),
    "mcv_csr_sddmm_create": (_I, [_LL, _LL, _LL, _LL, _P, _P, _P, _P, _P, _P, _P]),
    "mcv_gemm_create": (_I, [_I, _LL, _LL, _LL, _P, _P, _I, _P, _P, _P]),
    "mcv_run": (_I, [_P]),
    "mcv_workspace_bytes": (ctypes.c_size_t, [_P]),
    "mcv_destroy": (_I, [_P]),
    "mcv_last_error": (ctypes.c_char_p, []),
}
_vlib = None


def lib():
    """The loaded comparator library; raises when it has not been built."""
    global _vlib
    if _vlib is None:
        if not os.path.exists(VENDOR_PATH):
            raise RuntimeError(f"{VENDOR_PATH} is missing: run `python -m magicube_amd.build`")
        handle = ctypes.CDLL(VENDOR_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _vlib = handle
    return _vlib


def exported_symbols():
    return tuple(_SIGS)


class VendorError(RuntimeError):
    def __init__(self, status: int, where: str):
        detail = lib().mcv_last_error().decode(errors="replace")
        super().__init__(f"{where}: status {status} ({detail})")
        self.status = status


# ---- format conversion (host, numpy) --------------------------------------

def scalar_csr(csr: VectorCSR):
    """The vector-sparse pattern as a scalar CSR of m_vec * V rows.

    Returns (row_ptr [m + 1], col_ind [nnz_vec * V], gather) with the scalar
    values ``lanes[gather]`` for vector-major lanes (lane j * V + v holds row v
    of nonzero vector j, as in spmm_benchmark.cpp:89-118 of the reference).
    """
    V = csr.V
    ro = np.asarray(csr.row_offsets, dtype=np.int64)
    ci = np.asarray(csr.col_indices, dtype=np.int32)
    counts = np.diff(ro)
    row_len = np.repeat(counts, V)
    row_ptr = np.zeros(csr.m + 1, dtype=np.int64)
    np.cumsum(row_len, out=row_ptr[1:])
    # scalar row r = i * V + v takes slots ro[i] .. ro[i+1]-1 in order
    vec_row = np.repeat(np.arange(csr.m_vec), counts * V)
    r_of = np.repeat(np.arange(csr.m), row_len)
    pos = np.arange(row_ptr[-1]) - row_ptr[r_of]
    slot = ro[vec_row] + pos
    v = r_of % V
    gather = slot * V + v
    return row_ptr.astype(np.int32), ci[slot].astype(np.int32), gather.astype(np.int64)


def bsr_blocks(csr: VectorCSR, lanes: np.ndarray):
    """The vector-sparse operand in BSR with V x V blocks (row-major in block).

    A vector at column c lands in block column c // V, block column c % V;
    vectors of one vector row in the same aligned group share a block.
    Returns (row_ptr [m_vec + 1], col_ind [nnzb], values [nnzb * V * V]).
    """
    V = csr.V
    ro = np.asarray(csr.row_offsets, dtype=np.int64)
    ci = np.asarray(csr.col_indices, dtype=np.int64)
    if csr.k % V:
        raise ValueError("BSR needs k to be a multiple of V")
    vec_row = np.repeat(np.arange(csr.m_vec), np.diff(ro))
    key = vec_row * (csr.k // V) + ci // V
    uniq, block_of = np.unique(key, return_inverse=True)
    nnzb = uniq.size
    brow = uniq // (csr.k // V)
    row_ptr = np.zeros(csr.m_vec + 1, dtype=np.int64)
    np.add.at(row_ptr, brow + 1, 1)
    np.cumsum(row_ptr, out=row_ptr)
    vals = np.zeros((nnzb, V, V), dtype=np.asarray(lanes).dtype)
    lv = np.asarray(lanes).reshape(-1, V)
    vals[block_of, :, ci % V] = lv
    return row_ptr.astype(np.int32), (uniq % (csr.k // V)).astype(np.int32), vals.reshape(-1)


def uniform_blocks(m_vec: int, k: int, V: int, density: float, seed: int):
    """A BSR pattern with V x V blocks, the same number of blocks in every
    block row (int(density * k / V) + 1, distinct columns, ascending), drawn
    uniformly: the structure of the reference's Blocked-ELL baseline
    (baselines/spmm_benchmark.cpp:268-279, bm_test_utils.h:47-67).
    Returns (row_ptr [m_vec + 1], col_ind [m_vec * per_row])."""
    kb = k // V
    per_row = min(int(density * kb) + 1, kb)
    rng = np.random.default_rng(seed)
    keys = rng.random((m_vec, kb))
    cols = np.sort(np.argpartition(keys, per_row - 1, axis=1)[:, :per_row], axis=1)
    row_ptr = (np.arange(m_vec + 1, dtype=np.int64) * per_row).astype(np.int32)
    return row_ptr, cols.reshape(-1).astype(np.int32)


def dense_lhs(csr: VectorCSR, lanes: np.ndarray) -> np.ndarray:
    """The vector-sparse operand as a dense m x k row-major array."""
    V = csr.V
    out = np.zeros((csr.m_vec, V, csr.k), dtype=np.asarray(lanes).dtype)
    ro = np.asarray(csr.row_offsets, dtype=np.int64)
    vec_row = np.repeat(np.arange(csr.m_vec), np.diff(ro))
    out[vec_row, :, np.asarray(csr.col_indices, dtype=np.int64)] = np.asarray(lanes).reshape(-1, V)
    return out.reshape(csr.m, csr.k)


def vector_to_scalar_index(csr: VectorCSR) -> np.ndarray:
    """For each vector-major lane j * V + v (nonzero vector j, row v), its
    position in the scalar CSR of ``scalar_csr``: a scalar-CSR result array
    ``x`` read as ``x[vector_to_scalar_index(csr)]`` is in vector-major order."""
    _, _, gather = scalar_csr(csr)
    idx = np.empty_like(gather)
    idx[gather] = np.arange(gather.size)
    return idx


# ---- plans -----------------------------------------------------------------

def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream_handle(stream):
    import torch
    s = torch.cuda.current_stream() if stream is None else stream
    return ctypes.c_void_p(s.cuda_stream)


class Plan:
    """A created vendor plan; keeps its operand tensors alive."""

    def __init__(self, handle: ctypes.c_void_p, name: str, keep):
        self._h = handle
        self.name = name
        self._keep = keep

    def run(self):
        if self._h is None:
            raise RuntimeError(f"{self.name}: plan is closed")
        st = lib().mcv_run(self._h)
        if st != 0:
            raise VendorError(st, f"{self.name} run")

    @property
    def workspace_bytes(self) -> int:
        return int(lib().mcv_workspace_bytes(self._h)) if self._h is not None else 0

    def close(self):
        if self._h is not None:
            lib().mcv_destroy(self._h)
            self._h = None
            self._keep = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _create(fn_name: str, name: str, keep, *args):
    h = ctypes.c_void_p()
    st = getattr(lib(), fn_name)(*args, ctypes.byref(h))
    if st != 0:
        raise VendorError(st, name)
    return Plan(h, name, keep)


def csr_spmm(dtype: int, alg: str, m: int, k: int, n: int, row_ptr, col_ind, values, b, c, stream=None) -> Plan:
    """rocSPARSE CSR SpMM plan: c[m, n] = A[m, k] @ b[k, n] (device tensors)."""
    nnz = int(col_ind.numel())
    return _create("mcv_csr_spmm_create", f"rocsparse csr spmm ({alg})", (row_ptr, col_ind, values, b, c),
                   dtype, CSR_ALGS[alg], m, k, n, nnz, _ptr(row_ptr), _ptr(col_ind), _ptr(values), _ptr(b),
                   _ptr(c), _stream_handle(stream))


def bsr_spmm(dtype: int, mb: int, kb: int, n: int, block_dim: int, row_ptr, col_ind, values, b, c,
             stream=None) -> Plan:
    """rocSPARSE BSR SpMM plan: c[mb * bd, n] = A @ b[kb * bd, n]."""
    nnzb = int(col_ind.numel())
    return _create("mcv_bsr_spmm_create", "rocsparse bsr spmm", (row_ptr, col_ind, values, b, c), dtype, mb, kb, n,
                   nnzb, block_dim, _ptr(row_ptr), _ptr(col_ind), _ptr(values), _ptr(b), _ptr(c),
                   _stream_handle(stream))


def csr_sddmm(m: int, n: int, k: int, row_ptr, col_ind, lhs, rhs, out_values, stream=None) -> Plan:
    """rocSPARSE SDDMM plan (fp16 operands, fp32 values): out[j] = lhs[r] . rhs[c]."""
    nnz = int(col_ind.numel())
    return _create("mcv_csr_sddmm_create", "rocsparse sddmm", (row_ptr, col_ind, lhs, rhs, out_values), m, n, k,
                   nnz, _ptr(row_ptr), _ptr(col_ind), _ptr(lhs), _ptr(rhs), _ptr(out_values),
                   _stream_handle(stream))


def gemm(dtype: int, m: int, k: int, n: int, a, b, c, b_transposed: bool = False, stream=None) -> Plan:
    """hipBLASLt GEMM plan: c[m, n] = a[m, k] @ B, with b = B (k x n) or B^T (n x k)."""
    return _create("mcv_gemm_create", f"hipblaslt gemm ({'i8' if dtype == I8 else 'f16'}, "
                   f"{'TN' if b_transposed else 'NN'})", (a, b, c), dtype, m, k, n, _ptr(a), _ptr(b),
                   int(bool(b_transposed)), _ptr(c), _stream_handle(stream))
