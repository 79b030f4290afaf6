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


@pytest.mark.parametrize("rounding", [attention.ROUND_TRUNC, attention.ROUND_NEAREST])
@pytest.mark.parametrize("qk_bits", [8, 4])
@pytest.mark.parametrize("V", [8, 2])
def test_kernel_quantized_q_equals_quantize_launch(V, qk_bits, rounding):
    """The fused kernel's own quantization of fp16 Q (``fuse_q``) gives the
    quantize launch's integers: S, P, P V and O bit for bit with and without
    it, also with Q read in place from the middle of a [L][192] buffer.  Q
    rows that are not 16-byte aligned, or a row stride that is not a
    multiple of 8 halves, keep the quantize launch."""
    mats, host = make_mats([192, 512, 1000], V=V, sparsity=0.8, seed=70 + V + qk_bits + rounding)
    strided, unaligned, odd = [], [], []
    for m in mats:
        buf = torch.zeros(m.L, 192, dtype=torch.float16, device="cuda")
        buf[:, 64:128] = m.q
        strided.append(attention.AttentionMatrix(buf[:, 64:128], m.k, m.v, m.mask))
        flat = torch.zeros(m.L * 64 + 4, dtype=torch.float16, device="cuda")
        flat[4:] = m.q.reshape(-1)
        unaligned.append(attention.AttentionMatrix(flat[4:].view(m.L, 64), m.k, m.v, m.mask))
        buf2 = torch.zeros(m.L, 196, dtype=torch.float16, device="cuda")
        buf2[:, :64] = m.q
        odd.append(attention.AttentionMatrix(buf2[:, :64], m.k, m.v, m.mask))
    kw = dict(qk_bits=qk_bits, V=V, p_mode=attention.P_ROW_SCALED, p_scale=127.0, debug=True, rounding=rounding)
    ref = attention.AttentionBatch(mats, fuse_q=False, **kw)
    kw["scales"] = [d.scales for d in ref.mats]
    batches = [attention.AttentionBatch(x, **kw) for x in (mats, strided, unaligned, odd)]
    assert not ref.fuse_q and [b.fuse_q for b in batches] == [True, True, False, False]
    assert batches[0].n_first == len(mats) and ref.n_first == 2 * len(mats)     # K jobs only, or Q and K
    assert (batches[0].all_jobs is not None) == (qk_bits == 8)                    # int8 V joins int8 K
    print(f"kernel-quantized Q: V={V} qk_bits={qk_bits} rounding={rounding}", flush=True)
    for b in [ref] + batches:
        b.forward()
    torch.cuda.synchronize()
    want = _debug_outputs(ref)
    for b in batches:
        for d, w, g in zip(b.mats, want, _debug_outputs(b)):
            real = np.concatenate([np.arange(d.pairs[2 * i], d.pairs[2 * i + 1]) for i in range(d.m_vec)])
            idx = (real[:, None] * V + np.arange(V)[None, :]).ravel()
            assert np.array_equal(w[0][idx], g[0][idx]), "int32 logits differ"
            assert np.array_equal(w[1][: d.aligned_num * V], g[1][: d.aligned_num * V]), "int8 P differs"
            assert np.array_equal(w[2], g[2]), "int32 P V differs"
            assert np.array_equal(w[3].view(np.uint32), g[3].view(np.uint32)), "dequantized O differs"


@pytest.mark.parametrize("p_mode", [attention.P_ROW_SCALED, attention.P_REFERENCE, attention.P_REFERENCE_4BIT])
@pytest.mark.parametrize("qk_bits", [8, 4])
@pytest.mark.parametrize("V", [8, 4, 2])
def test_production_kernel_equals_plain(qk_bits, p_mode, V):
    """The fused kernel without debug outputs (the one callers run) gives the
    plain pipeline's fp32 O bit for bit (4-bit P runs with V in the 4-bit
    range: packed 4-bit V with 4-bit Q/K, one value per byte with int8 Q/K)."""
    mats, host = make_mats([128, 704, 1024], V=V, sparsity=0.8, seed=3 * V + qk_bits + p_mode)
    p_scale = 127.0 if p_mode == attention.P_ROW_SCALED else 32.0
    v4 = p_mode == attention.P_REFERENCE_4BIT
    b = attention.AttentionBatch(mats, qk_bits=qk_bits, V=V, p_mode=p_mode, p_scale=p_scale,
                                 v_range4=v4 and qk_bits == 8, v_bits=4 if v4 and qk_bits == 4 else 8)
    print(f"production fused vs plain V={V} qk_bits={qk_bits} p_mode={p_mode}", flush=True)
    b.forward_plain()
    torch.cuda.synchronize()
    plain = [d.o.cpu().numpy().copy() for d in b.mats]
    b.forward()
    torch.cuda.synchronize()
    for d, po in zip(b.mats, plain):
        assert np.array_equal(po.view(np.uint32), d.o.cpu().numpy().view(np.uint32))


@pytest.mark.parametrize("p_mode", [attention.P_ROW_SCALED, attention.P_REFERENCE, attention.P_REFERENCE_4BIT])
@pytest.mark.parametrize("V", [8, 4, 2])
def test_packed_4bit_v_equals_byte_v(V, p_mode):
    """Packed 4-bit V ([L][32] bytes, gathered 16 bits per lane and widened
    to V + 8 with 8 * sum(P) taken off the accumulators) gives the int32 S,
    P, P V and O of the same integers held one per byte, bit for bit; the
    packed operand is the nibble packing of the byte one; and the packed
    batch's one-op-per-launch pipeline equals its fused launch."""
    mats, host = make_mats([96, 640, 1000], V=V, sparsity=0.8, seed=60 + V + p_mode)
    p_scale = 127.0 if p_mode == attention.P_ROW_SCALED else 32.0
    kw = dict(qk_bits=4, V=V, p_mode=p_mode, p_scale=p_scale, debug=True)
    byte_v = attention.AttentionBatch(mats, v_range4=True, **kw)
    packed = attention.AttentionBatch(mats, v_bits=4, scales=[d.scales for d in byte_v.mats], **kw)
    assert packed.all_jobs is not None and byte_v.all_jobs is None     # V joins the Q/K launch
    print(f"packed 4-bit V vs byte V: V={V} p_mode={p_mode}", flush=True)
    byte_v.forward()
    packed.forward()
    torch.cuda.synchronize()
    for d8, d4, o8, o4 in zip(byte_v.mats, packed.mats, _debug_outputs(byte_v), _debug_outputs(packed)):
        v8 = d8.vq.cpu().numpy()[: d8.L * 64].view(np.int8).astype(np.int32)
        raw = d4.vq.cpu().numpy()[: d4.L * 32]
        nib = np.stack([raw & 15, raw >> 4], axis=1).ravel().astype(np.int32)
        assert np.array_equal(np.where(nib > 7, nib - 16, nib), v8), "packed V is not the nibble packing"
        for name, a, b in zip(("S", "P", "P V", "O"), o8, o4):
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), name
    fused = [d.o.cpu().numpy().copy() for d in packed.mats]
    packed.forward_plain()
    torch.cuda.synchronize()
    for d, fo in zip(packed.mats, fused):
        assert np.array_equal(d.o.cpu().numpy().view(np.uint32), fo.view(np.uint32)), "plain vs fused"


def test_packed_4bit_v_long_sequence():
    """The packed 4-bit V path at the serving trace's longest matrix (L = 8192,
    80 % sparse, about 1640 keys per row, V = 8) in the reference's 4-bit
    modes: S, P, P V and O equal those of byte-held V bit for bit."""
    mats, host = make_mats([8192], V=8, sparsity=0.8, seed=71)
    for p_mode in (attention.P_REFERENCE, attention.P_REFERENCE_4BIT):
        kw = dict(qk_bits=4, V=8, p_mode=p_mode, p_scale=32.0, scales=[(36.0, 36.0, 36.0)], debug=True)
        byte_v = attention.AttentionBatch(mats, v_range4=True, **kw)
        packed = attention.AttentionBatch(mats, v_bits=4, **kw)
        print(f"packed 4-bit V, L=8192, p_mode={p_mode}", flush=True)
        byte_v.forward()
        packed.forward()
        torch.cuda.synchronize()
        for name, a, b in zip(("S", "P", "P V", "O"), _debug_outputs(byte_v)[0], _debug_outputs(packed)[0]):
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), name


def _check_against_restatement(b, host, qk_bits, p_mode, p_scale, v_bits, V):
    """Every real slot of every vector row: int32 S exact, P within one step
    of the FP32 restatement (bitwise equal in most slots), P V from the
    GPU's own P exact, and in the reference modes the dequantized O equal to
    float32(P V) * o_scale bit for bit (q_softmax_kernel.cu:132-147,
    deq_spmm_kernel.cu:645-670)."""
    ref_mode = {attention.P_ROW_SCALED: 1, attention.P_REFERENCE: 0, attention.P_REFERENCE_4BIT: 2}[p_mode]
    p_max = 7 if p_mode == attention.P_REFERENCE_4BIT else 127
    for d, (q, k, v, mask) in zip(b.mats, host):
        ref = oatt.quantized_attention_ref(q, k, v, mask, d.scales, qk_bits=qk_bits, p_mode=ref_mode,
                                           p_scale=p_scale, v_bits=v_bits)
        s_out = d.s_out.cpu().numpy()
        p_out = d.p_out.cpu().numpy()
        o_int = d.o_int.cpu().numpy()
        o = d.o.cpu().numpy()
        worst, differ, total = 0, 0, 0
        for i in range(d.m_vec):
            s0, e0 = d.pairs[2 * i], d.pairs[2 * i + 1]
            slots = np.arange(s0, e0)
            S_gpu = s_out.reshape(-1, V)[slots].T
            assert np.array_equal(S_gpu, ref["S"][i]), f"S row {i}"
            # P in the SpMM A layout: slot s, row v at (s // 16) * 16V + 16 v + s % 16
            P_gpu = np.stack([p_out[(slots // 16) * 16 * V + 16 * vv + slots % 16] for vv in range(V)]).astype(np.int64)
            if slots.size:
                dif = np.abs(P_gpu - ref["P"][i])
                worst = max(worst, int(dif.max()))
                differ += int((dif != 0).sum())
                total += dif.size
                assert P_gpu.min() >= 0 and P_gpu.max() <= p_max, f"P row {i} outside [0, {p_max}]"
            # P V from the GPU's own P is exact
            cols = mask.col_indices[mask.row_offsets[i]:mask.row_offsets[i + 1]]
            oi = P_gpu @ ref["vq"][cols].astype(np.int64)
            assert np.array_equal(o_int[i * V:(i + 1) * V], oi), f"PV row {i}"
        assert worst <= 1, f"P differs from the FP32 restatement by {worst} steps"
        assert differ <= max(8, total // 100), f"{differ} of {total} P slots differ from the restatement"
        if p_mode != attention.P_ROW_SCALED:
            exp_o = o_int.astype(np.float32) * np.float32(d.o_scale)
            assert np.array_equal(o.view(np.uint32), exp_o.view(np.uint32)), "dequantized O"


P_MODES = [(attention.P_ROW_SCALED, 127.0, 8), (attention.P_REFERENCE, 32.0, 8),
           (attention.P_REFERENCE_4BIT, 32.0, 4)]


@pytest.mark.parametrize("p_mode,p_scale,v_bits", P_MODES)
@pytest.mark.parametrize("qk_bits", [8, 4])
def test_against_numpy_restatement(qk_bits, p_mode, p_scale, v_bits):
    V = 8
    mats, host = make_mats([512, 776], V=V, sparsity=0.9, seed=5)
    scales = None
    if p_mode != attention.P_ROW_SCALED:
        scales = [(36.0, 36.0, 36.0)] * len(mats)        # the reference layer's fixed scale (spattention.py:51-53)
    b = attention.AttentionBatch(mats, qk_bits=qk_bits, V=V, p_mode=p_mode, p_scale=p_scale, scales=scales,
                                 v_range4=v_bits == 4 and qk_bits == 8, v_bits=4 if v_bits == 4 and qk_bits == 4 else 8,
                                 debug=True)
    print(f"restatement qk_bits={qk_bits} p_mode={p_mode}", flush=True)
    b.forward()
    torch.cuda.synchronize()
    _check_against_restatement(b, host, qk_bits, p_mode, p_scale, v_bits, V)


@pytest.mark.parametrize("qk_bits", [8, 4])
def test_long_sequence_against_restatement_and_fp64(qk_bits):
    """The serving trace's longest matrix: L = 8192, 80 % sparse, V = 8:
    exact S, P within one step, exact P V, and the FP64 error bound."""
    V = 8
    mats, host = make_mats([8192], V=V, sparsity=0.8, seed=40 + qk_bits)
    b = attention.AttentionBatch(mats, qk_bits=qk_bits, V=V, debug=True)
    print(f"L=8192 restatement qk_bits={qk_bits}", flush=True)
    b.forward()
    torch.cuda.synchronize()
    _check_against_restatement(b, host, qk_bits, attention.P_ROW_SCALED, 127.0, 8, V)
    prod = attention.AttentionBatch(mats, qk_bits=qk_bits, V=V, scales=[d.scales for d in b.mats])
    o = prod.forward()[0].cpu().numpy()
    torch.cuda.synchronize()
    assert np.array_equal(o.view(np.uint32), b.mats[0].o.cpu().numpy().view(np.uint32))
    q, k, v, mask = host[0]
    err = oatt.error_vs_fp64(o, oatt.dense_attention_fp64(q, k, v, mask))
    print("L=8192 error vs FP64:", err, flush=True)
    # stated bounds at L = 8192, 80 % sparse (about 1640 keys per row): int8
    # Q/K/V (127 / max|x|) with row-scaled int8 P measures 0.036 on the host
    # restatement, int4 Q/K (7 / 2.5 rms) 0.165; more keys per row mean
    # smaller P entries, which lose more to int8 rounding than at L <= 4096
    bound = 0.045 if qk_bits == 8 else 0.2
    assert err["rel_fro"] < bound, err


def test_full_batch_8x16_fused_equals_plain():
    """The batched layer at its stated size: 8 samples x 16 heads,
    L ~ U[512, 4096] per sample, one 90 %-sparse V = 8 mask per (sample,
    head), in one persistent launch: the fused output equals the
    one-op-per-launch pipeline's bit for bit, and sampled matrices meet the
    FP64 bound."""
    V = 8
    rng = np.random.default_rng(77)
    Ls = [int(x) * 64 for x in rng.integers(512 // 64, 4096 // 64 + 1, size=8)]
    Ls[0] = 4096
    mats, host = [], []
    for s_i, L in enumerate(Ls):
        for h in range(16):
            sd = 500 + 1000 * s_i + h
            q, k, v = synthetic.attention_inputs(L, sd)
            mask = synthetic.attention_mask(L, V, 0.9, sd + 7)
            mats.append(attention.AttentionMatrix(*(torch.from_numpy(x).cuda() for x in (q, k, v)), mask))
            host.append((q, k, v, mask))
    b = attention.AttentionBatch(mats, qk_bits=8, V=V)
    print(f"8x16 batch, L = {Ls}", flush=True)
    b.forward_plain()
    torch.cuda.synchronize()
    plain = [d.o.clone() for d in b.mats]
    b.forward()
    torch.cuda.synchronize()
    for n, (p, d) in enumerate(zip(plain, b.mats)):
        assert torch.equal(p.view(torch.int32), d.o.view(torch.int32)), f"matrix {n}"
    for n in (0, 17, 127):
        q, k, v, mask = host[n]
        err = oatt.error_vs_fp64(b.mats[n].o.cpu().numpy(), oatt.dense_attention_fp64(q, k, v, mask))
        assert err["rel_fro"] < 0.035, (n, err)


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


def test_bquantization_on_a_side_stream():
    """Quantizing on another stream (its job descriptor, temporaries and
    output made on that stream) gives the default-stream result."""
    rng = np.random.default_rng(11)
    x = torch.from_numpy((rng.standard_normal((4, 1000, 64)) * 2).astype(np.float16)).cuda()
    ref = attention.bquantization(x, 8, 36.0)
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    for _ in range(3):
        got = attention.bquantization(x.transpose(0, 1).transpose(0, 1), 8, 36.0, stream=s)
        s.synchronize()
        assert torch.equal(got, ref)
        del got
