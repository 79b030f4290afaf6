"""Host logic of the serving loop that needs no GPU."""
import math

from magicube_amd import attention, serving


def test_percentile_matches_linear_interpolation():
    vals = sorted([5.0, 1.0, 3.0, 2.0, 4.0])
    assert serving._percentile(vals, 50) == 3.0
    assert serving._percentile(vals, 0) == 1.0 and serving._percentile(vals, 100) == 5.0
    assert math.isclose(serving._percentile(vals, 99), 4.96)


def test_working_set_counts_quantized_operands_and_output():
    # int8 q, k, v (L*64 each) + fp32 o (L*64*4)
    assert attention.working_set_bytes(1024, 8) == 1024 * 64 * 7
    # int4 q, k halve
    assert attention.working_set_bytes(1024, 4) == 1024 * 64 * 6


def _cpu_mats():
    import torch
    from magicube_amd import synthetic
    mats = []
    for n, L in enumerate((256, 520, 64)):
        q, k, v = (torch.from_numpy(x) for x in synthetic.attention_inputs(L, n))
        mats.append(attention.AttentionMatrix(q, k, v, synthetic.attention_mask(L, 8, 0.9, n)))
    scales = [(127.0 / 3.1, 127.0 / 2.9, 127.0 / 4.2 + n) for n in range(3)]
    return mats, scales


def test_resident_block_descriptors_match_per_matrix_path():
    """The vectorized block builds the same launch descriptors (fields and
    operand sizes) as the per-matrix path."""
    import numpy as np
    mats, scales = _cpu_mats()
    for bits in (8, 4):
        blk = attention.ResidentBlock(mats, scales, bits, 8, 127.0, device="cpu")
        devs = [attention.make_resident(m, sc, bits, 8, 127.0, "cpu") for m, sc in zip(mats, scales)]
        ref = np.frombuffer(b"".join(d.desc_bytes() for d in devs), dtype=attention._ATTN_DT)
        for f in ("L", "m_vec", "row_offsets", "col_indices", "alpha", "o_scale", "o_int", "p_out", "s_out"):
            if f in ("row_offsets", "col_indices"):
                continue            # the block uploads its own copy of a host mask
            assert np.array_equal(blk.desc[f], ref[f]), f
        assert not blk.desc["q16"].any()       # blocks launch with Q quantized by the quantize launch
        # operand regions: q, k sized L*64*bits/8, v L*64, o L*64*4, inside the arena, disjoint
        base = blk.arena.data_ptr()
        spans = []
        for i, d in enumerate(devs):
            for f, size in (("q", d.qq.numel()), ("k", d.kq.numel()), ("v", d.vq.numel()), ("o", d.o.numel() * 4)):
                start = int(blk.desc[f][i]) - base
                spans.append((start, start + size))
        spans.sort()
        assert spans[0][0] >= 0 and spans[-1][1] <= blk.arena.numel()
        assert all(a[1] <= b[0] for a, b in zip(spans, spans[1:]))
        jobs = np.frombuffer(b"".join(d.quant_job_bytes()[0] for d in devs), dtype=attention._JOB_DT)
        assert np.array_equal(blk.qk_jobs["count"], jobs["count"])
        assert np.array_equal(blk.qk_jobs["scale"], jobs["scale"])
        assert np.array_equal(blk.qk_jobs["x"], jobs["x"])
        outs = blk.outputs()
        assert [tuple(o.shape) for o in outs] == [(m.L, 64) for m in mats]
        assert [o.data_ptr() for o in outs] == [int(x) for x in blk.desc["o"]]


def test_block_batch_device_bytes_decode_to_valid_operands():
    """Decodes the descriptor and quantize-job buffers a block batch uploads,
    in the C record layout, and checks every pointer the kernels follow lands
    inside a live tensor of the right size (regression: concatenating padded
    records with numpy once repacked the jobs without their tail padding)."""
    import numpy as np
    import torch
    from magicube_amd import synthetic
    gen = torch.Generator().manual_seed(3)
    blocks, objs = [], []
    for rid, (bits, Ls) in enumerate(((8, (256, 512)), (8, (320,)), (4, (448, 64)))):
        mats = []
        for L in Ls:
            q, k, v = (torch.randn((L, 64), generator=gen).half() for _ in range(3))
            mats.append(attention.AttentionMatrix(q, k, v, synthetic.attention_mask_device(L, 8, 0.85, gen, "cpu")))
        blk = attention.ResidentBlock(mats, [(30.0, 31.0, 29.0 + rid)] * len(mats), bits, 8, 127.0, device="cpu")
        blocks.append((bits, blk))
        objs += [blk.arena] + [t for m in mats for t in (m.q, m.k, m.v, m.mask.row_offsets, m.mask.col_indices)]
    regs = [(t.data_ptr(), t.data_ptr() + t.numel() * t.element_size()) for t in objs]
    by_ptr = {t.data_ptr(): t for t in objs}

    def inside(p, n):
        return any(a <= p and p + n <= b for a, b in regs)

    for bits in (8, 4):
        bb = attention.BlockBatch([b for w, b in blocks if w == bits], device="cpu")
        raw = lambda t: t.numpy().tobytes()  # noqa: E731
        # 128-byte descriptors and 48-byte jobs (include/magicube_hip.h)
        assert attention._ATTN_DT.itemsize == 128 and attention._JOB_DT.itemsize == 48
        assert bb.desc.numel() == 128 * bb.n_mats and bb.qk_jobs.numel() == 2 * 48 * bb.n_mats
        assert bb.v_jobs.numel() == 48 * bb.n_mats
        desc = np.frombuffer(raw(bb.desc), dtype=attention._ATTN_DT)
        for j in np.frombuffer(raw(bb.qk_jobs), dtype=attention._JOB_DT).tolist():
            x, cnt, _, out, row_len, _ = j
            assert 0 < cnt <= bb.max_count and inside(x, 2 * cnt) and inside(out, cnt * bits // 8)
            assert row_len == 0          # block inputs are contiguous
        for x, cnt, _, out, row_len, _ in np.frombuffer(raw(bb.v_jobs), dtype=attention._JOB_DT).tolist():
            assert inside(x, 2 * cnt) and inside(out, cnt) and row_len == 0
        items = [it for it in bb.items.numpy().tolist() if it[0] >= 0]
        assert len(set(map(tuple, items))) == len(items) == sum(int(d["m_vec"]) for d in desc)
        assert bb.n_items == bb.items.shape[0]
        for e, row in items:
            d = desc[e]
            L = int(d["L"])
            assert 0 <= row < int(d["m_vec"])
            ro, ci = by_ptr[int(d["row_offsets"])], by_ptr[int(d["col_indices"])]
            s0, s1 = int(ro[2 * row]), int(ro[2 * row + 1])
            assert 0 <= s0 <= s1 <= ci.numel() and s0 % 16 == 0
            assert s1 == s0 or 0 <= int(ci[s0:s1].min()) and int(ci[s0:s1].max()) < L
            assert inside(int(d["o"]) + row * 8 * 256, 8 * 256) and int(d["o16"]) == 0
            for f, size in (("q", L * 8 * bits), ("k", L * 8 * bits), ("v", L * 64)):
                assert inside(int(d[f]), size)


def test_xcd_item_order_covers_every_row_once_and_keeps_matrices_on_one_xcd():
    import numpy as np
    m_vecs = [5, 0, 37, 12, 64, 3, 9, 21, 18, 40, 1]
    works = [10 * m + 7 * e for e, m in enumerate(m_vecs)]
    order = attention.xcd_item_order(m_vecs, works, n_xcd=8, wpb=4)
    real = order[order[:, 0] >= 0]
    exp = sorted((e, r) for e, m in enumerate(m_vecs) for r in range(m))
    assert sorted(map(tuple, real.tolist())) == exp
    assert order.shape[0] % 32 == 0 and (order[order[:, 0] < 0] == -1).all()
    xcd_of = {}
    for pos, (e, _) in enumerate(order.tolist()):
        if e >= 0:
            x = (pos // 4) % 8
            assert xcd_of.setdefault(e, x) == x, "a matrix spans two XCDs"
    assert attention.xcd_item_order([0, 0], [0, 0]).shape == (1, 2)


def test_record_dtypes_keep_c_layout_under_promotion():
    """The descriptor and job record dtypes keep the C structs' padding when
    numpy promotes them: concatenated or stacked record arrays still have
    the 128- and 48-byte strides the kernels read (the job struct has 4
    bytes of padding after its scale: a packed 44-byte stride would hand the
    kernel shifted pointers)."""
    import numpy as np
    for dt, size in ((attention._ATTN_DT, 128), (attention._JOB_DT, 48)):
        key = "L" if "L" in dt.names else "count"
        a = np.zeros(6, dtype=dt)
        a[key] = np.arange(6)
        for joined in (np.concatenate([a[:4].reshape(-1, 2)[:, 1], a[4:]]), np.stack([a[0], a[5]])):
            assert joined.dtype.itemsize == size
            raw = np.frombuffer(joined.tobytes(), dtype=dt)
            assert raw[key].tolist() == joined[key].tolist()
