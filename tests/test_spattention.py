"""Host-side parts of the sparse attention module (CPU): the reference's
mask builder restated, the reference-mask adapter's checks and the module's
argument checks."""
import numpy as np
import pytest
import torch

from magicube_amd import spattention


@pytest.mark.parametrize("align", [16, 32])
def test_static_random_mask_aligned_layout(align):
    m, n, s = 64, 512, 0.9
    cols, shuf, pairs, order, an = spattention.static_random_mask_aligned(m, n, s, align, seed=3, device="cpu")
    cols, shuf, pairs, order = (t.numpy() for t in (cols, shuf, pairs, order))
    starts, ends = pairs[0::2], pairs[1::2]
    counts = ends - starts
    assert counts.sum() == round((1 - s) * m * n)                 # exact density, as scipy.sparse.random
    assert (starts % align == 0).all() and an % align == 0 and cols.size == an
    assert np.array_equal(starts[1:], starts[:-1] + (counts[:-1] + align - 1) // align * align)
    for i in range(m):
        row = cols[starts[i]:ends[i]]
        assert (np.diff(row) > 0).all() and row.min(initial=0) >= 0 and row.max(initial=0) < n
        assert (cols[ends[i]:starts[i] + (counts[i] + align - 1) // align * align] == -1).all()
    g = shuf.reshape(-1, 8)
    assert np.array_equal(g[:, :4], cols.reshape(-1, 8)[:, 0::2]) and np.array_equal(g[:, 4:], cols.reshape(-1, 8)[:, 1::2])
    assert (np.diff(counts[order]) >= 0).all()                    # ascending length, as np.argsort


def test_mask_from_reference_checks():
    cols, _, pairs, order, an = spattention.static_random_mask_aligned(32, 256, 0.8, 16, seed=1, device="cpu")
    mk = spattention.mask_from_reference(order, pairs, cols, seq_len=256, vec_length=8)
    assert mk.m_vec == 32 and mk.aligned_num == an and mk.nnz_vec == int((pairs[1::2] - pairs[0::2]).sum())
    assert torch.equal(mk.row_nnz, pairs[1::2] - pairs[0::2])
    with pytest.raises(ValueError):
        spattention.mask_from_reference(order, pairs, cols, seq_len=512, vec_length=8)
    bad = pairs.clone()
    bad[2] += 1                                                   # a start off the 16-slot grid
    with pytest.raises(ValueError):
        spattention.mask_from_reference(order, bad, cols, seq_len=256, vec_length=8)


def test_module_argument_checks():
    mha = spattention.SpMultiheadAttention(embed_dim=128, num_heads=2)
    x = torch.zeros(64, 1, 128)
    cols, _, pairs, order, _ = spattention.static_random_mask_aligned(8, 64, 0.5, 16, seed=0, device="cpu")
    kw = dict(row_indices=order, row_offsets=pairs, column_indices=cols, vec_length=8)
    with pytest.raises(NotImplementedError):
        mha(x, x, x, need_weights=True, **kw)
    with pytest.raises(NotImplementedError):
        mha(x, x, x, need_weights=False, lhs_pre=4, rhs_pre=8, **kw)       # not one of the reference's pairs
    with pytest.raises(NotImplementedError):
        mha(x, x, x, need_weights=False, key_padding_mask=torch.zeros(1, 64, dtype=torch.bool), **kw)
    with pytest.raises(ValueError):
        spattention.SpMultiheadAttention(embed_dim=96, num_heads=2)    # head dimension 48


def test_plan_reads_projection_in_place_and_writes_heads_into_output():
    """The module's launch records point each (sample, head) at its slices of
    the [L][batch][3E] projection (row stride batch * 3E) and of the
    [L][batch][E] output (row stride batch * E), so no permute copy runs."""
    import numpy as np
    from magicube_amd import attention
    L, bsz, H = 128, 2, 3
    E = 64 * H
    cols, _, pairs, order, _ = spattention.static_random_mask_aligned(L // 8, L, 0.9, 16, seed=1, device="cpu")
    mask = spattention.mask_from_reference(order, pairs, cols, L, 8)
    plan = spattention._Plan(L, bsz, H, [mask] * bsz, 8, 8, "cpu")
    desc = np.frombuffer(plan.batch.desc.numpy().tobytes(), dtype=attention._ATTN_DT)
    qk = np.frombuffer(plan.batch.qk_jobs.numpy().tobytes(), dtype=attention._JOB_DT).reshape(-1, 2)
    vj = np.frombuffer(plan.batch.v_jobs.numpy().tobytes(), dtype=attention._JOB_DT)
    for b in range(bsz):
        for h in range(H):
            i = b * H + h
            assert int(desc["o"][i]) == 0 and int(desc["o16_stride"][i]) == bsz * E
            assert int(desc["o16"][i]) - plan.o.data_ptr() == 2 * (b * E + 64 * h)
            for s, job in enumerate((qk[i, 0], qk[i, 1], vj[i])):
                assert int(job["row_len"]) == 64 and int(job["row_stride"]) == bsz * 3 * E
                assert int(job["count"]) == L * 64
                assert int(job["x"]) - plan.qkv.data_ptr() == 2 * (b * 3 * E + s * E + 64 * h)


def test_make_resident_rejects_unaligned_fp16_outputs():
    import torch
    from magicube_amd import attention
    L = 64
    cols, _, pairs, order, _ = spattention.static_random_mask_aligned(L // 8, L, 0.9, 16, seed=2, device="cpu")
    mask = spattention.mask_from_reference(order, pairs, cols, L, 8)
    q = torch.zeros((L, 64), dtype=torch.float16)
    m = attention.AttentionMatrix(q, q, q, mask)
    bad = torch.zeros((L, 66), dtype=torch.float16)[:, :64]          # row stride 66: not a multiple of 4
    with pytest.raises(ValueError):
        attention.make_resident(m, (1.0, 1.0, 1.0), 8, 8, 32.0, "cpu", out=bad)
    ok = torch.zeros((L, 128), dtype=torch.float16)[:, 64:]
    d = attention.make_resident(m, (1.0, 1.0, 1.0), 8, 8, 32.0, "cpu", out=ok)
    assert d.o.data_ptr() == ok.data_ptr()


def test_make_resident_rejects_too_long_sequences(monkeypatch):
    import torch
    from magicube_amd import attention
    L = 64
    cols, _, pairs, order, _ = spattention.static_random_mask_aligned(L // 8, L, 0.9, 16, seed=3, device="cpu")
    mask = spattention.mask_from_reference(order, pairs, cols, L, 8)
    q = torch.zeros((L, 64), dtype=torch.float16)
    m = attention.AttentionMatrix(q, q, q, mask)
    assert attention.MAX_SEQ_LEN == 1 << 24
    monkeypatch.setattr(attention, "MAX_SEQ_LEN", L - 8)
    with pytest.raises(ValueError, match="sequence length"):
        attention.make_resident(m, (1.0, 1.0, 1.0), 8, 8, 32.0, "cpu")
    monkeypatch.setattr(attention, "MAX_SEQ_LEN", L)
    attention.make_resident(m, (1.0, 1.0, 1.0), 8, 8, 32.0, "cpu")


def test_bquantization_rejects_shapes_it_cannot_pack():
    import torch
    from magicube_amd import attention
    with pytest.raises(ValueError):
        attention.bquantization(torch.zeros(4099, dtype=torch.float16), 8, 1.0)   # 4099 bytes in 1024 words
    with pytest.raises(ValueError):
        attention.bquantization(torch.zeros(2, 12, dtype=torch.float16), 4, 1.0)  # 12 nibbles per row
    with pytest.raises(ValueError):
        attention.bquantization(torch.zeros(2, 64, dtype=torch.float16), 16, 1.0)


def test_spmm_and_sddmm_mirrors_reject_short_buffers():
    import torch
    from magicube_amd import sddmm, spmm
    z = torch.zeros(16, dtype=torch.int32)
    with pytest.raises(ValueError, match="output_matrix"):
        spmm.wmma_spmm(8, 8, 4, 8, 64, 64, z, z, z, z, torch.zeros(64 * 64 // 4, dtype=torch.int32),
                       torch.zeros(4 * 8 * 64 - 1, dtype=torch.int32))
    with pytest.raises(ValueError, match="rhs_matrix"):
        spmm.wmma_spmm(8, 4, 4, 8, 64, 64, z, z, z, z, torch.zeros(64 * 64 // 8 - 1, dtype=torch.int32),
                       torch.zeros(4 * 8 * 64, dtype=torch.int32))
    with pytest.raises(ValueError, match="lhs_matrix"):
        sddmm.wmma_sddmm(8, 4, 64, 64, z, z, z, torch.zeros(4 * 8 * 64 // 4 - 1, dtype=torch.int32),
                         torch.zeros(64 * 64 // 4, dtype=torch.int32), z, 8)


def test_encoder_embeds_transposed_tokens_contiguously():
    """The encoder's [L][batch][E] input is embedding(x.t()) * sqrt(E): the
    reference's transpose(embedding(x) * sqrt(E), 0, 1) bit for bit
    (end_to_end.py:160-166), laid out contiguously so the first block's
    LayerNorms take the fused kernel; the dense encoder's forward equals the
    transposed formulation's."""
    import math
    import torch
    torch.manual_seed(0)
    model = spattention.TransformerEncoder(vocab_size=37, embed_dim=64, num_heads=1, seq_len=32, mlp_dim=128,
                                           num_class=5, layers=1).eval()
    x = torch.randint(0, 37, (3, 32))
    ours = model.encoder(x.t()) * math.sqrt(64)
    ref = torch.transpose(model.encoder(x) * math.sqrt(64), 0, 1)
    assert ours.is_contiguous() and not ref.is_contiguous()
    assert torch.equal(ours, ref)
    with torch.no_grad():
        got = model(x)
        out = model.pos_encoder(ref)
        for blk in model.blocks:
            out = blk(out)
        want = model.linear(model.layer_norm(out))
    assert torch.allclose(got, want, rtol=0, atol=1e-6)
