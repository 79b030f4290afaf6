"""Quantized sparse attention on the GPU: quantizer exactness, the fused
batched launch against the one-op-per-launch pipeline (bitwise), the numpy
restatement (exact integers, P within one step), the FP64 dense-attention
error, alone-vs-mixed identity and hipGraph replays."""
import numpy as np
import pytest
import torch

from magicube_amd import attention, formats, synthetic
from oracle import attention as oatt

pytestmark = pytest.mark.gpu


def make_mats(Ls, V=8, sparsity=0.9, seed=0):
    mats, host = [], []
    for n, L in enumerate(Ls):
        q, k, v = synthetic.attention_inputs(L, seed + 17 * n)
        mask = synthetic.attention_mask(L, V, sparsity, seed + 31 * n)
        mats.append(attention.AttentionMatrix(*(torch.from_numpy(x).cuda() for x in (q, k, v)), mask))
        host.append((q, k, v, mask))
    return mats, host


def test_quantize_matches_reference():
    rng = np.random.default_rng(0)
    x = (rng.standard_normal(5000) * 3).astype(np.float16)
    xt = torch.from_numpy(x).cuda()
    for bits, scale in ((8, 36.0), (8, 127.0 / 9.0), (4, 2.5)):
        for rounding in (attention.ROUND_TRUNC, attention.ROUND_NEAREST):
            got = attention.bquantization(xt.view(1, -1, 8), bits, scale, rounding)
            torch.cuda.synchronize()
            exp = oatt.quantize_ref(x, scale, bits, rounding)
            lanes = formats.unpack_lanes(got.cpu().numpy().view(np.uint8), bits, x.size, signed=True)
            assert np.array_equal(lanes, exp), (bits, rounding)


def _debug_outputs(batch):
    return [(d.s_out.cpu().numpy(), d.p_out.cpu().numpy(), d.o_int.cpu().numpy(), d.o.cpu().numpy())
            for d in batch.mats]


@pytest.mark.parametrize("rounding", [attention.ROUND_TRUNC, attention.ROUND_NEAREST])
def test_quantize_4bit_range_in_int8(rounding):
    """MC_QUANT_RANGE4: int8 output clamped to [-8, 7], equal to the 4-bit
    restatement and to the 4-bit packed kernel's nibbles."""
    rng = np.random.default_rng(5)
    x = (rng.standard_normal(4104) * 3).astype(np.float16)          # rows of 8 values
    xt = torch.from_numpy(x).cuda().view(1, -1, 8)
    got = attention.bquantization(xt, 8, 2.5, rounding=rounding | attention.QUANT_RANGE4)
    assert got.numel() * 4 == x.size
    exp = oatt.quantize_ref(x, 2.5, 4, rounding)
    assert np.array_equal(got.cpu().numpy().view(np.int8).astype(np.int32).ravel(), exp)
    assert exp.min() == -8 and exp.max() == 7
    packed = attention.bquantization(xt, 4, 2.5, rounding=rounding).cpu().numpy().view(np.uint32).ravel()
    assert packed.size * 8 == x.size
    nib = np.stack([(packed >> (4 * e)) & 15 for e in range(8)], axis=1).ravel().astype(np.int32)
    assert np.array_equal(np.where(nib > 7, nib - 16, nib), exp)


@pytest.mark.parametrize("p_mode", [attention.P_ROW_SCALED, attention.P_REFERENCE])
@pytest.mark.parametrize("qk_bits", [8, 4])
@pytest.mark.parametrize("V", [8, 4, 2])
def test_fused_equals_plain(qk_bits, p_mode, V):
    mats, host = make_mats([256, 520, 1000], V=V, sparsity=0.85, seed=V + qk_bits)
    p_scale = 127.0 if p_mode == attention.P_ROW_SCALED else 32.0
    b = attention.AttentionBatch(mats, qk_bits=qk_bits, V=V, p_mode=p_mode, p_scale=p_scale, debug=True)
    print(f"fused vs plain V={V} qk_bits={qk_bits} p_mode={p_mode}", flush=True)
    b.forward_plain()
    torch.cuda.synchronize()
    plain = [(d.work["s"].cpu().numpy(), d.work["p"].cpu().numpy(), d.work["o_int"].cpu().numpy(),
              d.o.cpu().numpy().copy()) for d in b.mats]
    b.forward()
    torch.cuda.synchronize()
    fused = _debug_outputs(b)
    for d, (ps, pp, poi, po), (fs, fp, foi, fo) in zip(b.mats, plain, fused):
        real = np.concatenate([np.arange(d.pairs[2 * i], d.pairs[2 * i + 1]) for i in range(d.m_vec)])
        idx = (real[:, None] * V + np.arange(V)[None, :]).ravel()
        assert np.array_equal(ps[idx], fs[idx]), "int32 logits differ"
        assert np.array_equal(pp[: d.aligned_num * V], fp[: d.aligned_num * V]), "int8 P differs"
        assert np.array_equal(poi, foi), "int32 P V differs"
        assert np.array_equal(po.view(np.uint32), fo.view(np.uint32)), "dequantized O differs"


@pytest.mark.parametrize("p_mode", [attention.P_ROW_SCALED, attention.P_REFERENCE, attention.P_REFERENCE_4BIT])
@pytest.mark.parametrize("qk_bits", [8, 4])
@pytest.mark.parametrize("V", [8, 4, 2])
def test_production_kernel_equals_plain(qk_bits, p_mode, V):
    """The fused kernel without debug outputs (the one callers run) gives the
    plain pipeline's fp32 O bit for bit (4-bit P runs with V in the 4-bit range)."""
    mats, host = make_mats([128, 704, 1024], V=V, sparsity=0.8, seed=3 * V + qk_bits + p_mode)
    p_scale = 127.0 if p_mode == attention.P_ROW_SCALED else 32.0
    b = attention.AttentionBatch(mats, qk_bits=qk_bits, V=V, p_mode=p_mode, p_scale=p_scale,
                                 v_range4=p_mode == attention.P_REFERENCE_4BIT)
    print(f"production fused vs plain V={V} qk_bits={qk_bits} p_mode={p_mode}", flush=True)
    b.forward_plain()
    torch.cuda.synchronize()
    plain = [d.o.cpu().numpy().copy() for d in b.mats]
    b.forward()
    torch.cuda.synchronize()
    for d, po in zip(b.mats, plain):
        assert np.array_equal(po.view(np.uint32), d.o.cpu().numpy().view(np.uint32))


@pytest.mark.parametrize("qk_bits", [8, 4])
def test_against_numpy_restatement(qk_bits):
    V = 8
    mats, host = make_mats([512, 776], V=V, sparsity=0.9, seed=5)
    b = attention.AttentionBatch(mats, qk_bits=qk_bits, V=V, debug=True)
    b.forward()
    torch.cuda.synchronize()
    for d, (q, k, v, mask) in zip(b.mats, host):
        ref = oatt.quantized_attention_ref(q, k, v, mask, d.scales, qk_bits=qk_bits, p_mode=1, p_scale=127.0)
        s_out = d.s_out.cpu().numpy()
        p_out = d.p_out.cpu().numpy()
        o_int = d.o_int.cpu().numpy()
        worst = 0
        for i in range(d.m_vec):
            s0, e0 = d.pairs[2 * i], d.pairs[2 * i + 1]
            slots = np.arange(s0, e0)
            S_gpu = s_out.reshape(-1, V)[slots].T
            assert np.array_equal(S_gpu, ref["S"][i]), f"S row {i}"
            # P in the SpMM A layout: slot s, row v at (s // 16) * 16V + 16 v + s % 16
            P_gpu = np.stack([p_out[(slots // 16) * 16 * V + 16 * vv + slots % 16] for vv in range(V)]).astype(np.int64)
            worst = max(worst, int(np.abs(P_gpu - ref["P"][i]).max()) if slots.size else 0)
            # P V from the GPU's own P is exact
            cols = mask.col_indices[mask.row_offsets[i]:mask.row_offsets[i + 1]]
            oi = P_gpu @ ref["vq"][cols].astype(np.int64)
            assert np.array_equal(o_int[i * V:(i + 1) * V], oi), f"PV row {i}"
        assert worst <= 1, f"P differs from the FP32 restatement by {worst} steps"


def test_error_vs_fp64_oracle():
    V = 8
    mats, host = make_mats([1024, 2048], V=V, sparsity=0.9, seed=9)
    b = attention.AttentionBatch(mats, qk_bits=8, V=V, p_mode=attention.P_ROW_SCALED, p_scale=127.0)
    outs = [o.cpu().numpy() for o in b.forward()]
    torch.cuda.synchronize()
    errs = []
    for o, (q, k, v, mask) in zip(outs, host):
        ref = oatt.dense_attention_fp64(q, k, v, mask)
        errs.append(oatt.error_vs_fp64(o, ref))
    print("error vs FP64 (row-scaled int8 P, rounded):", errs, flush=True)
    # stated bound for int8 Q/K/V (per-tensor 127/max|x|) and row-scaled int8 P
    assert all(e["rel_fro"] < 0.035 for e in errs), errs


def test_alone_equals_in_the_mix():
    V = 8
    mats, _ = make_mats([256, 1200, 640, 2048, 512, 888], V=V, sparsity=0.9, seed=13)
    mix = attention.AttentionBatch(mats, qk_bits=8, V=V, debug=True)
    mix.forward(grid_waves=1024)
    torch.cuda.synchronize()
    mixed = _debug_outputs(mix)
    for n, m in enumerate(mats):
        alone = attention.AttentionBatch([m], qk_bits=8, V=V, scales=[mix.mats[n].scales], debug=True)
        alone.forward()
        torch.cuda.synchronize()
        (s1, p1, oi1, o1), = _debug_outputs(alone)
        s2, p2, oi2, o2 = mixed[n]
        assert np.array_equal(oi1, oi2) and np.array_equal(o1.view(np.uint32), o2.view(np.uint32)), n
        assert np.array_equal(p1, p2), n


def test_graph_replays_match_eager():
    V = 8
    mats, _ = make_mats([512, 1536, 800], V=V, sparsity=0.9, seed=21)
    b = attention.AttentionBatch(mats, qk_bits=8, V=V)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        eager = [o.clone() for o in b.forward(stream=s)]
    torch.cuda.synchronize()
    b.capture(s)
    for r in range(3):
        for o in b.mats:
            o.o.fill_(float("nan"))          # the replay must rewrite every output
        torch.cuda.synchronize()
        b.replay(s)
        torch.cuda.synchronize()
        x = torch.randn(1000, device="cuda").sum()   # unrelated eager work between replays
        for e, d in zip(eager, b.mats):
            assert torch.equal(e.view(torch.int32), d.o.view(torch.int32)), f"replay {r}"
        del x
