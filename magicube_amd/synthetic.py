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
