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
