"""Seeded synthetic inputs for the benchmark and test workloads.

The reference's sparsity patterns come from the DLMC ``.smtx`` files, which
are not available here; these generators produce the vector-sparse shapes the
workloads are stated on (uniform per-block masks, Poisson per-block nonzero
counts) and full-range integer operands.
"""
from __future__ import annotations

import numpy as np

from .formats import VectorCSR


def _rows_to_csr(m_vec: int, k: int, V: int, counts: np.ndarray, rng: np.random.Generator) -> VectorCSR:
    counts = np.clip(np.asarray(counts, dtype=np.int64), 0, k)
    offsets = np.zeros(m_vec + 1, dtype=np.int64)
    np.cumsum(counts, out=offsets[1:])
    cols = np.empty(int(offsets[-1]), dtype=np.int32)
    for i in range(m_vec):
        c = int(counts[i])
        if c == 0:
            continue
        if 4 * c < k:
            # sparse rows: draw until c distinct columns (cheap when c << k)
            pick = np.unique(rng.integers(0, k, size=c + c // 4 + 8))
            while pick.size < c:
                pick = np.unique(np.concatenate([pick, rng.integers(0, k, size=c)]))
            pick = rng.permutation(pick)[:c]
        else:
            pick = rng.permutation(k)[:c]
        cols[offsets[i]:offsets[i + 1]] = np.sort(pick)
    return VectorCSR(m_vec, k, V, offsets.astype(np.int32), cols)


def uniform_block_mask(m: int, k: int, V: int, sparsity: float, seed: int) -> VectorCSR:
    """Each V-row block keeps round((1 - sparsity) * k) column vectors, uniform without replacement."""
    rng = np.random.default_rng(seed)
    m_vec = m // V
    keep = int(round((1.0 - sparsity) * k))
    return _rows_to_csr(m_vec, k, V, np.full(m_vec, keep), rng)


def poisson_block_mask(m: int, k: int, V: int, sparsity: float, seed: int) -> VectorCSR:
    """Per V-row block nonzero count ~ Poisson((1 - sparsity) * k), clamped to [0, k]."""
    rng = np.random.default_rng(seed)
    m_vec = m // V
    counts = rng.poisson((1.0 - sparsity) * k, size=m_vec)
    return _rows_to_csr(m_vec, k, V, counts, rng)


def random_lanes(count: int, bits: int, signed: bool, seed: int) -> np.ndarray:
    """Uniform full-range integer lanes (int32 holding the lane value)."""
    rng = np.random.default_rng(seed)
    if signed:
        lo, hi = -(1 << (bits - 1)), (1 << (bits - 1))
    else:
        lo, hi = 0, 1 << bits
    return rng.integers(lo, hi, size=count, dtype=np.int64).astype(np.int32)


def attention_mask(L: int, V: int, sparsity: float, seed: int) -> VectorCSR:
    """L x L vector-sparse mask with V-row blocks: each (block, column) kept
    with probability 1 - sparsity (seeded Bernoulli, vectorized per block)."""
    rng = np.random.default_rng(seed)
    m_vec = L // V
    keep = rng.random((m_vec, L), dtype=np.float32) < np.float32(1.0 - sparsity)
    rows, cols = np.nonzero(keep)
    offsets = np.zeros(m_vec + 1, dtype=np.int64)
    np.cumsum(np.bincount(rows, minlength=m_vec), out=offsets[1:])
    return VectorCSR(m_vec, L, V, offsets.astype(np.int32), cols.astype(np.int32))


def attention_inputs(L: int, seed: int, d: int = 64):
    """q, k, v ~ N(0, 1) as float16 arrays of shape [L, d]."""
    rng = np.random.default_rng(seed)
    return tuple(rng.standard_normal((L, d), dtype=np.float32).astype(np.float16) for _ in range(3))


def attention_mask_device(L: int, V: int, sparsity: float, generator, device="cuda"):
    """Device-generated L x L vector-sparse mask in the kernels' format
    (a :class:`magicube_amd.attention.ResidentMask`): each (V-row block,
    column) kept with probability 1 - sparsity, rows padded to 16 slots with
    -1.  Used where host generation would dominate (long sequences, many
    requests)."""
    import torch
    from .attention import ResidentMask
    m_vec = L // V
    keep = torch.rand((m_vec, L), generator=generator, device=device) < (1.0 - sparsity)
    counts = keep.sum(dim=1)
    aligned = (counts + 15) // 16 * 16
    a_start = torch.cumsum(aligned, 0) - aligned
    u_start = torch.cumsum(counts, 0) - counts
    rows, cols = keep.nonzero(as_tuple=True)          # row-major, columns ascending
    total = int(aligned.sum())
    col_idx = torch.full((max(total, 16),), -1, dtype=torch.int32, device=device)
    col_idx[a_start[rows] + (torch.arange(rows.numel(), device=device) - u_start[rows])] = cols.to(torch.int32)
    pairs = torch.stack([a_start, a_start + counts], dim=1).reshape(-1).to(torch.int32)
    order = torch.argsort(counts, descending=True, stable=True).to(torch.int32)
    nnz = int(rows.numel())
    return ResidentMask(L=L, V=V, m_vec=m_vec, nnz_vec=nnz, aligned_num=total,
                        row_offsets=pairs if m_vec else torch.zeros(2, dtype=torch.int32, device=device),
                        col_indices=col_idx,
                        row_indices=order if m_vec else torch.zeros(1, dtype=torch.int32, device=device),
                        row_nnz=counts.to(torch.int32) if m_vec else torch.zeros(1, dtype=torch.int32, device=device))
