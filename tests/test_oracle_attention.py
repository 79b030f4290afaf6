"""Pins the attention oracles (oracle/attention.py) on the CPU: quantization
against a scalar loop, the FP64 dense attention against torch's float64
softmax, exact integers by brute force, closed-form single-column rows, the
quantized layer's distance from FP64, and planted mistakes that the FP64
comparison must expose."""
import math

import numpy as np
import pytest
import torch

from magicube_amd import formats, synthetic
from oracle import attention as oatt


def _scalar_quant(x, scale, bits, rounding):
    lo, hi = (-128.0, 127.0) if bits == 8 else (-8.0, 7.0)
    f = float(np.float32(np.float32(x) * np.float32(scale)))
    f = min(max(f, lo), hi)
    return int(math.trunc(f)) if rounding == 0 else int(round(f))   # round(): half to even


@pytest.mark.parametrize("bits", [8, 4])
@pytest.mark.parametrize("rounding", [0, 1])
def test_quantize_ref_matches_scalar_loop(bits, rounding):
    edges = [0.0, 0.5, -0.5, 1.5, -1.5, 2.5, -2.5, 0.49, 126.6, 127.5, -128.7, 300.0, -300.0, 7.5, -8.5, 6.9]
    rng = np.random.default_rng(1)
    x = np.concatenate([np.array(edges), rng.standard_normal(200) * 4]).astype(np.float16)
    for scale in (1.0, 36.0 / 9.0, 0.73):
        got = oatt.quantize_ref(x, scale, bits, rounding)
        exp = [_scalar_quant(v, scale, bits, rounding) for v in x]
        assert got.tolist() == exp


def _tiny(L=64, V=8, sparsity=0.7, seed=0, empty_row=True):
    q, k, v = synthetic.attention_inputs(L, seed)
    mask = synthetic.attention_mask(L, V, sparsity, seed + 1)
    if empty_row:                                   # vector row 1 keeps no column
        ro, ci = mask.row_offsets.copy(), mask.col_indices
        n1 = ro[2] - ro[1]
        ci = np.concatenate([ci[:ro[1]], ci[ro[2]:]])
        ro[2:] -= n1
        mask = formats.VectorCSR(mask.m_vec, L, V, ro, ci)
    return q, k, v, mask


def test_dense_fp64_matches_torch_masked_softmax():
    q, k, v, mask = _tiny()
    L, V = q.shape[0], mask.V
    keep = np.zeros((L, L), bool)
    for i in range(mask.m_vec):
        keep[i * V:(i + 1) * V, mask.col_indices[mask.row_offsets[i]:mask.row_offsets[i + 1]]] = True
    s = torch.from_numpy(q).double() @ torch.from_numpy(k).double().T / 8.0
    p = torch.softmax(s.masked_fill(~torch.from_numpy(keep), float("-inf")), dim=1).nan_to_num(0.0)
    exp = (p @ torch.from_numpy(v).double()).numpy()
    got = oatt.dense_attention_fp64(q, k, v, mask)
    assert np.allclose(got, exp, rtol=0, atol=1e-12)
    assert (got[8:16] == 0).all()


def test_single_column_rows_are_closed_form():
    L, V = 32, 8
    q, k, v = synthetic.attention_inputs(L, 3)
    cols = np.array([5, 17, 0, 31], np.int32)
    mask = formats.VectorCSR(4, L, V, np.arange(5, dtype=np.int32), cols)
    o = oatt.dense_attention_fp64(q, k, v, mask)
    for i, c in enumerate(cols):                    # softmax over one column = 1
        assert np.array_equal(o[i * V:(i + 1) * V], np.repeat(v[c:c + 1].astype(np.float64), V, 0))
    scales = (30.0, 31.0, 32.0)
    r = oatt.quantized_attention_ref(q, k, v, mask, scales, p_mode=1, p_scale=127.0)
    for i, c in enumerate(cols):
        assert (r["P"][i] == 127).all()
        assert np.array_equal(r["o_int"][i * V:(i + 1) * V], 127 * np.repeat(r["vq"][c:c + 1], V, 0))
    r0 = oatt.quantized_attention_ref(q, k, v, mask, scales, p_mode=0, p_scale=32.0)
    assert all((P == 32).all() for P in r0["P"])


def test_quantized_integers_by_brute_force():
    q, k, v, mask = _tiny(L=32, sparsity=0.5, seed=4)
    V = mask.V
    r = oatt.quantized_attention_ref(q, k, v, mask, (20.0, 21.0, 22.0), qk_bits=8)
    qq, kq, vq = r["qq"].tolist(), r["kq"].tolist(), r["vq"].tolist()
    for i in range(mask.m_vec):
        cols = mask.col_indices[mask.row_offsets[i]:mask.row_offsets[i + 1]].tolist()
        for rr in range(V):
            for j, c in enumerate(cols):
                assert r["S"][i][rr, j] == sum(a * b for a, b in zip(qq[i * V + rr], kq[c]))
            for d in range(64):
                exp = sum(int(r["P"][i][rr, j]) * vq[c][d] for j, c in enumerate(cols))
                assert r["o_int"][i * V + rr, d] == exp


def test_quantized_layer_is_close_to_fp64():
    q, k, v = synthetic.attention_inputs(512, 7)
    mask = synthetic.attention_mask(512, 8, 0.9, 8)
    sc = tuple(127.0 / float(np.abs(t.astype(np.float32)).max()) for t in (q, k, v))
    ref = oatt.dense_attention_fp64(q, k, v, mask)
    good = oatt.error_vs_fp64(oatt.quantized_attention_ref(q, k, v, mask, sc)["o"], ref)
    assert good["rel_fro"] < 0.035, good


@pytest.mark.parametrize("mistake", ["swapped_qk", "missing_sqrt_d", "shifted_mask", "sign_flip_v"])
def test_planted_mistakes_move_far_from_fp64(mistake):
    q, k, v = synthetic.attention_inputs(512, 7)
    mask = synthetic.attention_mask(512, 8, 0.9, 8)
    sc = [127.0 / float(np.abs(t.astype(np.float32)).max()) for t in (q, k, v)]
    ref = oatt.dense_attention_fp64(q, k, v, mask)
    qq, kk, vv, mm = q, k, v, mask
    if mistake == "swapped_qk":
        qq, kk = k, q
    elif mistake == "missing_sqrt_d":
        # same quantized integers, alpha 8x too large: 1 / sqrt(d) dropped
        qq = (q.astype(np.float32) * 8).astype(np.float16)
        sc[0] /= 8.0
    elif mistake == "shifted_mask":
        mm = formats.VectorCSR(mask.m_vec, mask.k, mask.V, mask.row_offsets, (mask.col_indices + 1) % mask.k)
    elif mistake == "sign_flip_v":
        vv = (-v.astype(np.float32)).astype(np.float16)
    bad = oatt.error_vs_fp64(oatt.quantized_attention_ref(qq, kk, vv, mm, tuple(sc))["o"], ref)
    assert bad["rel_fro"] > 0.2, (mistake, bad)


def test_4bit_p_and_v_range_modes():
    """p_mode 2 is the reference requantization clamped at 7 (a single-column
    row gives P = min(32, 7) = 7; elsewhere P = min(P_reference, 7) since
    trunc(min(x, 7)) = min(trunc(min(x, 127)), 7) for x >= 0), and v_bits 4
    clamps V to [-8, 7] before truncation."""
    L, V = 32, 8
    q, k, v = synthetic.attention_inputs(L, 3)
    cols = np.array([5, 17, 0, 31], np.int32)
    mask = formats.VectorCSR(4, L, V, np.arange(5, dtype=np.int32), cols)
    r = oatt.quantized_attention_ref(q, k, v, mask, (30.0, 31.0, 32.0), p_mode=2, p_scale=32.0)
    assert all((P == 7).all() for P in r["P"])
    q, k, v, mask = _tiny(L=64, sparsity=0.5, seed=9)
    sc = (36.0, 36.0, 36.0 * 8)
    r0 = oatt.quantized_attention_ref(q, k, v, mask, sc, p_mode=0, p_scale=32.0)
    r2 = oatt.quantized_attention_ref(q, k, v, mask, sc, p_mode=2, p_scale=32.0, v_bits=4)
    for P0, P2 in zip(r0["P"], r2["P"]):
        assert np.array_equal(P2, np.minimum(P0, 7))
    assert np.array_equal(r2["vq"], np.clip(r0["vq"], -8, 7))
    assert r2["vq"].min() == -8 and r2["vq"].max() == 7       # the clamp is exercised
    assert any((P0 > 7).any() for P0 in r0["P"])


# --------------------------------------------------------------------------
# the reference's per-op numerics (bsddmm, q_bcsr_softmax, bspmm)
# --------------------------------------------------------------------------

def _aligned_mask(m_vec, n, density, align, seed):
    """Reference-style aligned mask: (start, end) pairs, -1 padded columns."""
    rng = np.random.default_rng(seed)
    pairs, cols, pos = [], [], 0
    for i in range(m_vec):
        c = np.sort(rng.choice(n, size=int(rng.binomial(n, density)), replace=False))
        if i == 1:
            c = c[:0]                                   # an empty row
        pairs += [pos, pos + c.size]
        padded = -(-c.size // align) * align
        cols += list(c) + [-1] * (padded - c.size)
        pos += padded
    return np.array(pairs, np.int32), np.array(cols, np.int32)


@pytest.mark.parametrize("bits", [16, 8, 4])
@pytest.mark.parametrize("V", [2, 4, 8])
def test_reference_p_layout_round_trip(bits, V):
    rng = np.random.default_rng(bits * 10 + V)
    nnz = 64
    lo, hi = {16: (-32768, 32767), 8: (-128, 127), 4: (-8, 7)}[bits]
    P = rng.integers(lo, hi + 1, size=(nnz, V))
    raw = oatt.pack_p_reference(P, V, bits)
    assert raw.size == nnz * V * bits // 8
    assert np.array_equal(oatt.unpack_p_reference(raw, nnz, V, bits), P)


def test_reference_p_layout_known_positions():
    # byte offsets computed by hand from q_softmax_kernel.cu:253-268 for V = 2
    P = np.zeros((32, 2), np.int64)
    P[3, 1] = 5                                  # 8-bit: (3/16)*2*16 + 16*1 + 3 = 19
    raw = oatt.pack_p_reference(P, 2, 8)
    assert raw[19] == 5 and raw.sum() == 5
    P[:] = 0
    P[18, 0] = 0x0102                            # 16-bit: (18/16)*2*32 + 0 + 2 = 66, high byte +16
    raw = oatt.pack_p_reference(P, 2, 16)
    assert raw[66] == 2 and raw[82] == 1 and raw.sum() == 3
    P[:] = 0
    P[5, 1] = 3                                  # 4-bit: byte 16*1 + 5/2 = 18, odd slot -> high nibble
    P[4, 1] = 6
    raw = oatt.pack_p_reference(P, 2, 4)
    assert raw[18] == 0x36 and raw.sum() == 0x36


@pytest.mark.parametrize("bits", [16, 8, 4])
def test_ref_softmax_f16_closed_forms(bits):
    V, scale = 4, 32.0
    hi = {16: 32767, 8: 127, 4: 7}[bits]
    # row 0: one slot -> P = trunc(min(scale, hi)); row 1: three equal logits -> trunc(scale / 3)
    pairs = np.array([0, 1, 32, 35], np.int32)
    vals = np.zeros((1, 64 * V), np.float16)
    vals[0, 32 * V:35 * V] = np.float16(0.75)
    vals[0, :V] = np.float16(-2.5)
    P = oatt.ref_softmax_f16(vals, pairs, [0, 1], V, 0.125, scale, bits)
    assert (P[0, 0] == min(32, hi)).all()
    assert (P[0, 32:35] == min(10, hi)).all()               # trunc(32 / 3) = 10
    assert (P[0, 1:32] == 0).all() and (P[0, 35:] == 0).all()


def test_ref_softmax_f16_against_float64():
    V, scale, sqrt_dk = 8, 32.0, 0.125
    pairs, cols = _aligned_mask(12, 256, 0.2, 16, seed=3)
    rng = np.random.default_rng(4)
    vals = (rng.standard_normal((2, cols.size * V)) * 8).astype(np.float16)
    P = oatt.ref_softmax_f16(vals, pairs, list(range(12)), V, sqrt_dk, scale, 8)
    worst = 0
    for i in range(12):
        s0, e0 = pairs[2 * i], pairs[2 * i + 1]
        if e0 == s0:
            continue
        for b in range(2):
            s = vals[b].reshape(-1, V)[s0:e0].astype(np.float64) * sqrt_dk
            p = np.exp(s - s.max(axis=0)) / np.exp(s - s.max(axis=0)).sum(axis=0)
            worst = max(worst, int(np.abs(P[b, s0:e0] - np.trunc(p * scale)).max()))
    assert worst <= 1
    # a dropped 1/sqrt(d) is caught
    bad = oatt.ref_softmax_f16(vals, pairs, list(range(12)), V, 1.0, scale, 8)
    assert not np.array_equal(bad, P)


def test_ref_bsddmm_and_bspmm_against_dense():
    V, k, L = 4, 64, 96
    pairs, cols = _aligned_mask(L // V, L, 0.3, 16, seed=7)
    rng = np.random.default_rng(8)
    lhs = rng.integers(-128, 128, size=(2, L, k))
    rhs = rng.integers(-128, 128, size=(2, L, k))
    S = oatt.ref_bsddmm(lhs, rhs, pairs, cols, V, 1296.0).reshape(2, -1, V)
    for b in range(2):
        dense = lhs[b] @ rhs[b].T                                   # [L, L]
        for i in range(L // V):
            s0, e0 = pairs[2 * i], pairs[2 * i + 1]
            exp = (dense[i * V:(i + 1) * V][:, cols[s0:e0]].T.astype(np.float32) / np.float32(1296)).astype(np.float16)
            assert np.array_equal(S[b, s0:e0].view(np.uint16), exp.view(np.uint16))
            pad_end = pairs[2 * i + 2] if i + 1 < L // V else cols.size
            assert (S[b, e0:pad_end] == 0).all()
    # bspmm: the sparse P as a dense [L, L] matrix times V
    P = rng.integers(0, 128, size=(2, cols.size, V))
    vl = rng.integers(-128, 128, size=(2, L, 64))
    O = oatt.ref_bspmm(P, vl, pairs, cols, V, 1152.0)
    for b in range(2):
        D = np.zeros((L, L), np.int64)
        for i in range(L // V):
            s0, e0 = pairs[2 * i], pairs[2 * i + 1]
            D[i * V:(i + 1) * V, cols[s0:e0]] += P[b, s0:e0].T
        exp = ((D @ vl[b]).astype(np.float32) / np.float32(1152)).astype(np.float16)
        assert np.array_equal(O[b].view(np.uint16), exp.view(np.uint16))
    # a shifted column index is caught
    bad = oatt.ref_bspmm(P, vl, pairs, np.where(cols >= 0, (cols + 1) % L, -1), V, 1152.0)
    assert not np.array_equal(bad, O)
