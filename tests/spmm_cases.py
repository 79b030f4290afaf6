"""Shared SpMM case builder for the GPU tests (oracle used only as checker)."""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from magicube_amd import formats, synthetic


@dataclass
class SpmmCase:
    lhs_bits: int
    rhs_bits: int
    V: int
    signed: bool
    m_vec: int
    K: int
    N: int
    csr: formats.VectorCSR
    a: np.ndarray       # lane values, nnz_vec * V (in the mode's interpretation)
    b: np.ndarray       # lane values, K * N
    prep: formats.PreparedSpmm
    rhs_packed: np.ndarray  # int32 words, K * N * rhs_bits / 32


def make_case(lhs_bits, rhs_bits, V, signed, M, K, N, sparsity, seed, mask="poisson", a=None, b=None) -> SpmmCase:
    """Seeded case; ``a`` / ``b`` replace the generated lane values when given."""
    gen = synthetic.poisson_block_mask if mask == "poisson" else synthetic.uniform_block_mask
    csr = gen(M, K, V, sparsity, seed)
    if a is None:
        a = synthetic.random_lanes(csr.nnz_vec * V, lhs_bits, signed, seed + 1)
    if b is None:
        b = synthetic.random_lanes(K * N, rhs_bits, signed, seed + 2)
    prep = formats.prepare_spmm(csr, formats.pack_lanes(a, lhs_bits), lhs_bits, rhs_bits)
    rhs = formats.pack_lanes(b, rhs_bits)
    rhs = np.concatenate([rhs, np.zeros((-rhs.size) % 4, np.uint8)]).view(np.int32)
    return SpmmCase(lhs_bits, rhs_bits, V, signed, csr.m_vec, K, N, csr, a, b, prep, rhs)


def checksum_expected(case: SpmmCase, w: np.ndarray) -> np.ndarray:
    """(A (B w)) mod 2^32 per output row, exact in int64 for these ranges."""
    Bw = case.b.astype(np.int64).reshape(case.K, case.N) @ w.astype(np.int64)
    av = case.a.astype(np.int64).reshape(-1, case.V)
    ro, ci = case.csr.row_offsets, case.csr.col_indices
    contrib = av * Bw[ci][:, None]
    out = np.zeros((case.m_vec, case.V), dtype=np.int64)
    rows = np.repeat(np.arange(case.m_vec), np.diff(ro))
    np.add.at(out, rows, contrib)
    return (out.reshape(-1) & 0xFFFFFFFF).astype(np.uint32)


def checksum_got(C: np.ndarray, w: np.ndarray) -> np.ndarray:
    return ((C.astype(np.int64) @ w.astype(np.int64)) & 0xFFFFFFFF).astype(np.uint32)
