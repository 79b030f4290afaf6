"""Product host preparation (mc_format_*) against the oracle's restatement of
BmFN's preparation (SpMM/SpMM/spmm_benchmark.cpp:177-377), byte for byte."""
import numpy as np
import pytest

from magicube_amd import formats, synthetic
from tests.test_oracle import REF_TABLE, small_mask


@pytest.mark.parametrize("align", [8, 16, 32])
def test_align_rows_and_cols(orc, align):
    csr = small_mask(33, 200, 4, 0.2, seed=align)
    p_pairs, p_n = formats.align_rows(csr.row_offsets, align)
    o_pairs, o_n = orc.align_row_offsets(csr.row_offsets, align)
    assert p_n == o_n and np.array_equal(p_pairs, o_pairs)
    p_cols = formats.align_cols(csr.row_offsets, csr.col_indices, p_pairs, p_n, shuffle=False)
    o_cols = orc.align_col_indices(csr.row_offsets, csr.col_indices, o_pairs, o_n)
    assert np.array_equal(p_cols, o_cols)
    p_sh = formats.align_cols(csr.row_offsets, csr.col_indices, p_pairs, p_n, shuffle=True)
    assert np.array_equal(p_sh, orc.shuffle_col_indices(o_cols))


@pytest.mark.parametrize("V", [2, 4, 8])
@pytest.mark.parametrize("preA,preB", [(8, 8), (16, 8), (8, 4), (4, 4), (12, 8), (16, 16), (12, 4), (16, 4)])
def test_prepare_values_matches_reference(orc, V, preA, preB):
    m_vec, K = 24, 160
    tb, pa, cut, scale = REF_TABLE[(preA, preB, V)]
    csr = small_mask(m_vec, K, V, 0.25, seed=preA * 10 + V)
    nnz = csr.nnz_vec
    if preA == 12:
        # full 16-bit containers: both preparations keep bits 0..11 only
        packed = formats.pack_lanes(synthetic.random_lanes(nnz * V, 16, True, seed=5), 16)
    else:
        lanes = synthetic.random_lanes(nnz * V, preA, True, seed=5)
        packed = formats.pack_lanes(lanes, preA)
    align = formats.row_alignment(preA, preB)
    pairs, an = formats.align_rows(csr.row_offsets, align)
    mine = formats.prepare_values(preA, preB, V, csr.row_offsets, pairs, an, packed)
    words = np.concatenate([packed, np.zeros((-packed.size) % tb, np.uint8)])
    ref = orc.prepare_values(words.view({1: np.int8, 2: np.int16, 4: np.int32, 8: np.int64}[tb]), tb, scale,
                             csr.row_offsets, pairs, an, V, pa, cut, preB)
    assert mine.size == ref.size
    assert np.array_equal(mine, ref)


def test_sorted_rows(orc):
    csr = small_mask(64, 128, 8, 0.3, seed=9)
    mine = formats.sorted_rows(csr.row_offsets)
    assert np.array_equal(np.sort(mine), np.arange(64))
    lens = np.diff(csr.row_offsets)[mine]
    assert (np.diff(lens) <= 0).all()
    assert np.array_equal(mine, orc.sorted_row_swizzle(csr.row_offsets))


def test_pack_unpack_12_bit_containers():
    lanes = synthetic.random_lanes(777, 12, True, seed=12)
    packed = formats.pack_lanes(lanes, 12)
    assert packed.size == 2 * lanes.size
    assert (packed.view(np.uint16) < 4096).all()
    assert np.array_equal(formats.unpack_lanes(packed, 12, lanes.size, True), lanes)
    u = synthetic.random_lanes(777, 12, False, seed=13)
    assert np.array_equal(formats.unpack_lanes(formats.pack_lanes(u, 12), 12, u.size, False), u)


@pytest.mark.parametrize("bits", [4, 8, 16])
@pytest.mark.parametrize("signed", [True, False])
def test_pack_unpack_match_oracle(orc, bits, signed):
    lanes = synthetic.random_lanes(1001, bits, signed, seed=bits)
    mine = formats.pack_lanes(lanes, bits)
    assert np.array_equal(mine, orc.pack_lanes(lanes, bits))
    assert np.array_equal(formats.unpack_lanes(mine, bits, lanes.size, signed), lanes)
    assert np.array_equal(orc.unpack_lanes(mine, bits, lanes.size, signed), lanes)


def test_prepare_spmm_bundle():
    csr = synthetic.uniform_block_mask(512, 512, 8, 0.9, seed=1)
    assert np.all(np.diff(csr.row_offsets) == 51)
    vals = formats.pack_lanes(synthetic.random_lanes(csr.nnz_vec * 8, 8, True, 2), 8)
    prep = formats.prepare_spmm(csr, vals, 8, 8)
    assert prep.aligned_num == 64 * 64
    assert prep.values.size == prep.aligned_num * 8
    assert prep.col_indices.size == prep.aligned_num
    prep4 = formats.prepare_spmm(csr, vals, 8, 4)
    assert prep4.aligned_num % 32 == 0


def test_poisson_mask_statistics():
    csr = synthetic.poisson_block_mask(4096, 4096, 8, 0.9, seed=3)
    counts = np.diff(csr.row_offsets)
    assert csr.m_vec == 512
    assert abs(counts.mean() - 409.6) < 5
    for i in range(0, 512, 97):
        c = csr.col_indices[csr.row_offsets[i]:csr.row_offsets[i + 1]]
        assert np.all(np.diff(c) > 0) and c.min() >= 0 and c.max() < 4096
