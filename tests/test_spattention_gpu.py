"""The sparse multi-head attention module and the encoder on the GPU: the
module's attention equals the one-op-per-launch pipeline bitwise and the
numpy restatement of the reference's quantized attention within one P step;
the encoder is deterministic and replays in a HIP graph bitwise."""
import numpy as np
import pytest
import torch

import oracle.attention as oatt
from magicube_amd import formats, spattention

pytestmark = pytest.mark.gpu


def _layer(L, bsz, heads, rhs_pre, seed, lhs_pre=8):
    torch.manual_seed(seed)
    E = 64 * heads
    mha = spattention.SpMultiheadAttention(embed_dim=E, num_heads=heads).cuda().half().eval()
    align = 32 if rhs_pre == 4 else 16
    cols, _, pairs, order, _ = spattention.static_random_mask_aligned(L // 8, L, 0.9, align, seed=seed)
    x = (torch.randn(L, bsz, E, device="cuda") * 0.5).half()
    return mha, x, dict(row_indices=order, row_offsets=pairs, column_indices=cols, vec_length=8, rhs_pre=rhs_pre,
                        lhs_pre=lhs_pre)


PAIRS = [(8, 8), (16, 8), (8, 4), (4, 4)]


@pytest.mark.parametrize("pair", PAIRS, ids=lambda p: f"L{p[0]}R{p[1]}")
def test_module_equals_plain_pipeline_and_restatement(orc, pair):
    lhs_pre, rhs_pre = pair
    L, bsz, heads = 512, 2, 2
    mha, x, kw = _layer(L, bsz, heads, rhs_pre, seed=7, lhs_pre=lhs_pre)
    print(f"spattention L={L} bsz={bsz} heads={heads} lhs_pre={lhs_pre} rhs_pre={rhs_pre}", flush=True)
    with torch.no_grad():
        out, _ = mha(x, x, x, need_weights=False, **kw)
        torch.cuda.synchronize()
        plan = next(iter(mha._plans.values()))
        # the fused kernel writes each head's fp16 output into [L][batch][heads * 64]
        fused = plan.o.view(L, bsz, heads, 64).permute(1, 2, 0, 3).reshape(bsz * heads, L, 64).cpu().numpy()
        qkv = plan.qkv.view(L, bsz, 3, heads, 64).permute(2, 1, 3, 0, 4).reshape(3, bsz * heads, L, 64)
        stage = qkv.cpu().numpy()
        plain = [o.cpu().numpy() for o in plan.batch.forward_plain()]
        torch.cuda.synchronize()
    assert torch.isfinite(out).all()
    for i in range(bsz * heads):
        # the same fp32 values, rounded to fp16 to nearest even in the kernel
        assert np.array_equal(fused[i].view(np.uint16), plain[i].astype(np.float16).view(np.uint16)), f"matrix {i}"
    # the reference's formulas restated in numpy (static scales 36 / 32, truncation)
    pairs = kw["row_offsets"].cpu().numpy()
    cols = kw["column_indices"].cpu().numpy()
    counts = pairs[1::2] - pairs[0::2]
    ro = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    ci = np.concatenate([cols[pairs[2 * i]:pairs[2 * i + 1]] for i in range(L // 8)]).astype(np.int32)
    csr = formats.VectorCSR(L // 8, L, 8, ro, ci)
    for i in range(bsz * heads):
        ref = oatt.quantized_attention_ref(stage[0, i], stage[1, i], stage[2, i], csr, (36.0, 36.0, 36.0),
                                           qk_bits=rhs_pre, p_mode=2 if lhs_pre == 4 else 0, p_scale=32.0,
                                           rounding=0, v_bits=rhs_pre)
        # P agrees with the FP32 restatement to within one step (1/32 of a
        # probability) at a few slots; each such slot moves O by at most |v| / 32
        # slot moves O by at most |v| / 32; the fp16 store adds half an fp16 ulp
        vmax = np.abs(ref["vq"]).max() / 36.0
        got = fused[i].astype(np.float32)
        diff = np.abs(got - ref["o"])
        assert diff.max() <= 4 * vmax / 32.0 + 2.0 ** -11 * np.abs(ref["o"]).max() + 1e-6, f"matrix {i}"
        assert np.linalg.norm(got - ref["o"]) <= 1e-2 * np.linalg.norm(ref["o"]), f"matrix {i}"


def test_16bit_p_equals_8bit_p():
    """P <= 32 at the reference's scales, so lhs_pre 16 gives the lhs_pre 8 layer's output bit for bit."""
    L, bsz, heads = 256, 2, 2
    mha, x, kw = _layer(L, bsz, heads, 8, seed=13)
    print("spattention lhs_pre 16 vs 8", flush=True)
    with torch.no_grad():
        o8, _ = mha(x, x, x, need_weights=False, **kw)
        o16, _ = mha(x, x, x, need_weights=False, **dict(kw, lhs_pre=16))
        torch.cuda.synchronize()
    assert torch.equal(o8, o16)


def test_4bit_p_and_v_ranges():
    """lhs_pre 4 clamps P at 7; rhs_pre 4 keeps V in [-8, 7] and uses all of it."""
    L, bsz, heads = 256, 1, 2
    mha, x, kw = _layer(L, bsz, heads, 4, seed=17, lhs_pre=4)
    print("spattention lhs_pre 4 rhs_pre 4 ranges", flush=True)
    with torch.no_grad():
        mha(x * 8, x * 8, x * 8, need_weights=False, **kw)
        torch.cuda.synchronize()
        plan = next(iter(mha._plans.values()))
        vq = torch.cat([d.vq.view(torch.int8).flatten().cpu() for d in plan.batch.mats])
    assert int(vq.min()) == -8 and int(vq.max()) == 7


def test_module_output_projection_and_layout():
    L, bsz, heads = 256, 3, 2
    mha, x, kw = _layer(L, bsz, heads, 8, seed=11)
    with torch.no_grad():
        out, _ = mha(x, x, x, need_weights=False, **kw)
        plan = next(iter(mha._plans.values()))
        o = plan.o.clone()
        exp = torch.nn.functional.linear(o, mha.out_proj.weight, mha.out_proj.bias)
        # each (sample, head) lands at [:, sample, head * 64:(head + 1) * 64]
        plain = plan.batch.forward_plain()
        torch.cuda.synchronize()
        for b in range(bsz):
            for h in range(heads):
                assert torch.equal(o[:, b, 64 * h:64 * (h + 1)], plain[b * heads + h].half()), (b, h)
        # in-projection: q, k, v of (sample b, head h) are read from qkv in place
        qkv = torch.nn.functional.linear(x, mha.in_proj_weight, mha.in_proj_bias)
        assert torch.equal(plan.qkv, qkv)
    torch.cuda.synchronize()
    assert torch.equal(out, exp)


def test_encoder_deterministic_and_graph_replay():
    torch.manual_seed(0)
    model = spattention.SparseTransformerEncoder(vocab_size=257, embed_dim=256, num_heads=4, sparsity=0.9,
                                                 seq_len=1024, mlp_dim=1024, num_class=2, vec_length=8,
                                                 seed=5).cuda().half().eval()
    x = torch.randint(0, 257, (2, 1024), dtype=torch.int32, device="cuda")
    print("encoder seq 1024 bs 2", flush=True)
    s = torch.cuda.Stream()
    with torch.no_grad(), torch.cuda.stream(s):
        a = model(x)
        b = model(x)
        s.synchronize()
        assert torch.isfinite(a.float()).all() and torch.equal(a, b)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            c = model(x)
        for _ in range(3):
            g.replay()
            s.synchronize()
            assert torch.equal(c, a)


def _dense_p_reference(mha, x, kw, L, bsz, heads):
    """The plain pipeline's int8 P of every (sample, head) as dense [bsz, heads, L, L] / 32."""
    plan = next(iter(mha._plans.values()))
    plan.batch.forward_plain()
    torch.cuda.synchronize()
    dense = np.zeros((bsz, heads, L, L), np.float64)
    for n, d in enumerate(plan.batch.mats):
        p = d.work["p"].cpu().numpy().view(np.int8)
        pairs, cols = d.pairs, d.col_indices.cpu().numpy()
        for i in range(d.m_vec):
            for j in range(pairs[2 * i], pairs[2 * i + 1]):
                for v in range(8):
                    dense[n // heads, n % heads, i * 8 + v, cols[j]] = p[(j // 16) * 128 + 16 * v + j % 16] / 32.0
    return dense


def test_need_weights_returns_head_averaged_quantized_probabilities():
    L, bsz, heads = 256, 2, 2
    mha, x, kw = _layer(L, bsz, heads, 8, seed=17)
    print("spattention need_weights", flush=True)
    with torch.no_grad():
        out_w, w = mha(x, x, x, need_weights=True, **kw)
        out, none = mha(x, x, x, need_weights=False, **kw)
        torch.cuda.synchronize()
    assert none is None and w.shape == (bsz, L, L) and w.dtype == x.dtype
    assert torch.equal(out_w, out), "the weights path gives the same output"
    exp = _dense_p_reference(mha, x, kw, L, bsz, heads).mean(axis=1)
    got = w.float().cpu().numpy()
    assert np.allclose(got, exp, atol=2e-3, rtol=0), np.abs(got - exp).max()
    assert (got >= 0).all() and (got.sum(axis=2) <= 1.0 + 1e-3).all()


def test_key_padding_mask_removes_padded_keys():
    """A padding mask gives each sample the shared mask minus its padded key
    columns (checked against a hand-filtered mask row by row), padded keys
    get no weight, the fused output equals the plain pipeline on those
    masks, and an all-False padding mask changes nothing."""
    L, bsz, heads = 256, 2, 2
    mha, x, kw = _layer(L, bsz, heads, 8, seed=19)
    pad = torch.zeros(bsz, L, dtype=torch.bool, device="cuda")
    pad[0, 200:] = True
    pad[1, :32] = True
    print("spattention key_padding_mask", flush=True)
    with torch.no_grad():
        base, _ = mha(x, x, x, need_weights=False, **kw)
        none_padded, _ = mha(x, x, x, key_padding_mask=torch.zeros_like(pad), need_weights=False, **kw)
        out, w = mha(x, x, x, key_padding_mask=pad, need_weights=True, **kw)
        torch.cuda.synchronize()
    assert torch.equal(base, none_padded)
    assert not torch.equal(base, out)
    wn = w.float().cpu().numpy()
    assert (wn[0, :, 200:] == 0).all() and (wn[1, :, :32] == 0).all()
    plan = list(mha._plans.values())[-1]
    pairs, cols = kw["row_offsets"].cpu().numpy(), kw["column_indices"].cpu().numpy()
    for b in range(bsz):
        drop = pad[b].cpu().numpy()
        m = plan.masks[b]
        mp, mc = m.row_offsets.cpu().numpy(), m.col_indices.cpu().numpy()
        for i in range(L // 8):
            c = cols[pairs[2 * i]:pairs[2 * i + 1]]
            assert np.array_equal(np.sort(mc[mp[2 * i]:mp[2 * i + 1]]), np.sort(c[~drop[c]])), (b, i)
    with torch.no_grad():
        fused = plan.o.clone()
        plain = plan.batch.forward_plain()
        torch.cuda.synchronize()
    heads_out = fused.view(L, bsz, heads, 64)
    for n, o in enumerate(plain):
        assert np.array_equal(heads_out[:, n // heads, n % heads].cpu().numpy().view(np.uint16),
                              o.cpu().numpy().astype(np.float16).view(np.uint16)), f"matrix {n}"


def test_plan_cache_is_bounded_and_follows_mask_writes():
    L, bsz, heads = 128, 1, 1
    mha, x, kw = _layer(L, bsz, heads, 8, seed=23)
    with torch.no_grad():
        a, _ = mha(x, x, x, need_weights=False, **kw)
        # an in-place write to the caller's mask makes a new plan
        kw["column_indices"][kw["row_offsets"][0]] = (kw["column_indices"][kw["row_offsets"][0]] + 1) % L
        b, _ = mha(x, x, x, need_weights=False, **kw)
        torch.cuda.synchronize()
    assert len(mha._plans) == 2 and not torch.equal(a, b)
    for s in range(spattention.MAX_PLANS + 3):
        _, _, kws = _layer(L, bsz, heads, 8, seed=100 + s)
        with torch.no_grad():
            mha(x, x, x, need_weights=False, **kws)
    assert len(mha._plans) == spattention.MAX_PLANS
    mha.clear_plans()
    assert len(mha._plans) == 0
