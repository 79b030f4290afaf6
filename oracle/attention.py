"""Attention oracles — TEST INFRASTRUCTURE ONLY.

* :func:`dense_attention_fp64`: the FP64 reference the quantized layer states
  its error against: softmax(Q K^T / sqrt(d)) V in float64 on the fp16
  inputs, restricted to the vector-sparse mask (row r of vector row i attends
  to vector row i's columns; a row with no column yields zeros).
* :func:`quantized_attention_ref`: numpy restatement of the quantized layer
  (quantization_kernel.cu:46-74 for 8 bits, :13-43 for 4 bits, truncating,
  or rounding to nearest in the row-scaled layer;
  spattention.py:182-216; q_softmax_kernel.cu:132-147 for the reference
  requantization, :84-118 for its 4-bit P, p_mode 2) — exact integers for Qq/Kq/Vq and S; FP32 softmax in numpy's
  own summation order (the GPU's fixed order may differ by one rounding, so P
  is compared within one quantization step); P V and the dequantization given
  a P are exact.
"""
from __future__ import annotations

import math

import numpy as np

HEAD_DIM = 64


def _rows(csr):
    ro, ci = csr.row_offsets, csr.col_indices
    return [ci[ro[i]:ro[i + 1]] for i in range(csr.m_vec)]


def dense_attention_fp64(q16: np.ndarray, k16: np.ndarray, v16: np.ndarray, csr) -> np.ndarray:
    q = q16.astype(np.float64)
    k = k16.astype(np.float64)
    v = v16.astype(np.float64)
    V = csr.V
    out = np.zeros_like(q)
    scale = 1.0 / math.sqrt(q.shape[1])
    for i, cols in enumerate(_rows(csr)):
        if cols.size == 0:
            continue
        s = q[i * V:(i + 1) * V] @ k[cols].T * scale          # [V, nnz]
        s = s - s.max(axis=1, keepdims=True)
        p = np.exp(s)
        p /= p.sum(axis=1, keepdims=True)
        out[i * V:(i + 1) * V] = p @ v[cols]
    return out


def quantize_ref(x16: np.ndarray, scale: float, bits: int, rounding: int = 0) -> np.ndarray:
    """trunc (rounding 0) or round-half-even (1) of clamp(float32(x) * float32(scale))."""
    lo, hi = (-128.0, 127.0) if bits == 8 else (-8.0, 7.0)
    f = x16.astype(np.float32) * np.float32(scale)
    f = np.clip(f, np.float32(lo), np.float32(hi))
    return (np.rint(f) if rounding else np.trunc(f)).astype(np.int32)


def quantized_attention_ref(q16, k16, v16, csr, scales, qk_bits=8, p_mode=1, p_scale=127.0, rounding=None,
                            v_bits=8):
    """Returns dict(qq, kq, vq, S (list of [V, nnz] int64), P (list of [V, nnz] int),
    o_int [L, 64] int64 computed from this P, o float32).  p_mode 0: reference
    requantization (P clamped at 127), 1: row-scaled, 2: reference into 4-bit
    P (clamped at 7).  v_bits: the range V is quantized to (8 or 4)."""
    sq, sk, sv = scales
    if rounding is None:
        rounding = 1 if p_mode == 1 else 0
    qq = quantize_ref(q16, sq, qk_bits, rounding)
    kq = quantize_ref(k16, sk, qk_bits, rounding)
    vq = quantize_ref(v16, sv, v_bits, rounding)
    p_max = np.float32(7.0 if p_mode == 2 else 127.0)
    V = csr.V
    alpha = np.float32(1.0) / (np.float32(sq) * np.float32(sk) * np.float32(math.sqrt(HEAD_DIM)))
    o_scale = np.float32(1.0) / (np.float32(sv) * np.float32(p_scale))
    L = q16.shape[0]
    S_rows, P_rows = [], []
    o_int = np.zeros((L, HEAD_DIM), dtype=np.int64)
    o = np.zeros((L, HEAD_DIM), dtype=np.float32)
    for i, cols in enumerate(_rows(csr)):
        S = qq[i * V:(i + 1) * V].astype(np.int64) @ kq[cols].astype(np.int64).T
        S_rows.append(S)
        if cols.size == 0:
            P_rows.append(np.zeros((V, 0), np.int64))
            continue
        s = S.astype(np.float32) * alpha
        m = s.max(axis=1, keepdims=True)
        e = np.exp(s - m).astype(np.float32)
        l = e.sum(axis=1, dtype=np.float32, keepdims=True)
        inv_l = np.float32(1.0) / l
        if p_mode != 1:
            P = np.trunc(np.minimum(e * inv_l * np.float32(p_scale), p_max)).astype(np.int64)
        else:
            P = np.rint(np.minimum(e * np.float32(p_scale), p_max)).astype(np.int64)
        P_rows.append(P)
        oi = P @ vq[cols].astype(np.int64)
        o_int[i * V:(i + 1) * V] = oi
        f = oi.astype(np.float32) * o_scale
        if p_mode == 1:
            f = f * inv_l
        o[i * V:(i + 1) * V] = f
    return dict(qq=qq, kq=kq, vq=vq, S=S_rows, P=P_rows, o_int=o_int, o=o, alpha=alpha, o_scale=o_scale)


def error_vs_fp64(o: np.ndarray, o_ref: np.ndarray) -> dict:
    diff = o.astype(np.float64) - o_ref
    den = np.linalg.norm(o_ref)
    return {"max_abs": float(np.abs(diff).max()) if diff.size else 0.0,
            "rel_fro": float(np.linalg.norm(diff) / den) if den > 0 else 0.0}


# ---------------------------------------------------------------------------
# The reference's per-op numerics (the sptrans ops of
# end2end_eval/sparse_transformer_magicube/src/cuda).  Masks are the aligned
# vector-CSR: pairs = 2 * m_vec (start, end), cols = aligned column indices.
# ---------------------------------------------------------------------------

def ref_bsddmm(lhs: np.ndarray, rhs: np.ndarray, pairs, cols, V: int, scale: float) -> np.ndarray:
    """fp16 [batch, nnz * V]: half(float32(S) / float32(scale)) with S the
    exact int sum over k of lhs[b, i V + v, k] * rhs[b, col_j, k] at every real
    slot j of vector row i, vector-major [slot][v]; padding slots 0
    (deq_sddmm_kernel.cu:565-636, sddmm_utils/output_tile.h:37-43).
    lhs [batch, m, k] and rhs [batch, n, k] hold integer lanes."""
    batch = lhs.shape[0]
    nnz = len(cols)
    out = np.zeros((batch, nnz, V), dtype=np.float16)
    for i in range(len(pairs) // 2):
        s0, e0 = int(pairs[2 * i]), int(pairs[2 * i + 1])
        if e0 <= s0:
            continue
        c = np.asarray(cols[s0:e0], dtype=np.int64)
        for b in range(batch):
            S = lhs[b, i * V:(i + 1) * V].astype(np.int64) @ rhs[b, c].astype(np.int64).T     # [V, n]
            out[b, s0:e0] = (S.T.astype(np.float32) / np.float32(scale)).astype(np.float16)
    return out.reshape(batch, nnz * V)


def ref_softmax_f16(values: np.ndarray, pairs, rows, V: int, sqrt_dk: float, scale: float, bits: int) -> np.ndarray:
    """int P [batch, nnz, V] of the reference's quantized softmax
    (q_softmax_kernel.cu:180-271 with vsoftmax_16b/_8b/_4b :84-177):
    s = float32(half) * sqrt_dk; the running max of row v starts at -1e10
    for v = 0 and at 0 for v >= 1 (the register array is initialised as
    {-1e+10}, :209); P = trunc(clamp(exp(s - max) / sum * scale)) in the
    width's range; padding slots 0.  The sum runs in numpy's order (the
    reference reduces with CUB), so P may differ by one step from any
    fixed-order evaluation."""
    lo, hi = {16: (-32768.0, 32767.0), 8: (-128.0, 127.0), 4: (-8.0, 7.0)}[bits]
    batch = values.shape[0]
    nnz = values.shape[1] // V
    vals = values.reshape(batch, nnz, V)
    P = np.zeros((batch, nnz, V), dtype=np.int64)
    init = np.array([-1e10] + [0.0] * (V - 1), dtype=np.float32)
    for m in rows:
        s0, e0 = int(pairs[2 * m]), int(pairs[2 * m + 1])
        if e0 <= s0:
            continue
        for b in range(batch):
            s = vals[b, s0:e0].astype(np.float32) * np.float32(sqrt_dk)           # [nz, V]
            mx = np.maximum(s.max(axis=0), init)
            e = np.exp(s - mx).astype(np.float32)
            sm = e.sum(axis=0, dtype=np.float32)
            p = (e / sm).astype(np.float32) * np.float32(scale)
            P[b, s0:e0] = np.trunc(np.clip(p, lo, hi)).astype(np.int64)
    return P


def pack_p_reference(P: np.ndarray, V: int, bits: int) -> np.ndarray:
    """[nnz, V] P -> the reference's packed bytes (nnz * V * bits / 8):
    8-bit [(i/16)][v][16]; 16-bit [(i/16)][v][low 16][high 16]; 4-bit
    [(i/32)][v][16] with byte j = slot 2j (low nibble) | 2j+1 (high)
    (q_softmax_kernel.cu:253-268).  nnz must be a multiple of 16 (32 for
    4-bit)."""
    nnz = P.shape[0]
    p = P.astype(np.int64)
    if bits == 8:
        return (p.reshape(nnz // 16, 16, V).transpose(0, 2, 1) & 255).astype(np.uint8).ravel()
    if bits == 16:
        g = p.reshape(nnz // 16, 16, V).transpose(0, 2, 1) & 0xFFFF          # [G, V, 16]
        return np.stack([g & 255, g >> 8], axis=2).astype(np.uint8).ravel()   # [G, V, 2, 16]
    g = p.reshape(nnz // 32, 32, V).transpose(0, 2, 1) & 15                  # [G, V, 32]
    return (g[..., 0::2] | (g[..., 1::2] << 4)).astype(np.uint8).ravel()


def unpack_p_reference(raw: np.ndarray, nnz: int, V: int, bits: int) -> np.ndarray:
    """Inverse of :func:`pack_p_reference` (signed values)."""
    b = np.asarray(raw, dtype=np.uint8)
    if bits == 8:
        g = b[: nnz * V].reshape(nnz // 16, V, 16).astype(np.int8).astype(np.int64)
        return g.transpose(0, 2, 1).reshape(nnz, V)
    if bits == 16:
        g = b[: nnz * V * 2].reshape(nnz // 16, V, 2, 16).astype(np.int64)
        w = (g[:, :, 0] | (g[:, :, 1] << 8)).astype(np.uint16).astype(np.int16).astype(np.int64)
        return w.transpose(0, 2, 1).reshape(nnz, V)
    g = b[: nnz * V // 2].reshape(nnz // 32, V, 16).astype(np.int64)
    nib = np.stack([g & 15, g >> 4], axis=3).reshape(nnz // 32, V, 32)
    nib = np.where(nib > 7, nib - 16, nib)
    return nib.transpose(0, 2, 1).reshape(nnz, V)


def ref_bspmm(P: np.ndarray, v_lanes: np.ndarray, pairs, cols, V: int, scale: float) -> np.ndarray:
    """fp16 [batch, m_vec * V, n]: half(float32(O) / float32(scale)) with O the
    exact int sum of P[b, j, v] * v_lanes[b, col_j, :] over the real slots of
    each vector row (deq_spmm_kernel.cu:645-670,1287-1361)."""
    batch, _, n = v_lanes.shape
    m_vec = len(pairs) // 2
    out = np.zeros((batch, m_vec * V, n), dtype=np.float16)
    for i in range(m_vec):
        s0, e0 = int(pairs[2 * i]), int(pairs[2 * i + 1])
        c = np.asarray(cols[s0:e0], dtype=np.int64)
        for b in range(batch):
            O = P[b, s0:e0].astype(np.int64).T @ v_lanes[b, c].astype(np.int64)              # [V, n]
            out[b, i * V:(i + 1) * V] = (O.astype(np.float32) / np.float32(scale)).astype(np.float16)
    return out
