"""SDDMM parity on the GPU through the C ABI: bit-exact against the oracle
(restated Host_sddmm_integers, SDDMM/SDDMM/sddmm_benchmark.cpp:24-60, and
its plain signed/unsigned form)."""
import numpy as np
import pytest
import torch

from magicube_amd import formats, sddmm, synthetic

pytestmark = pytest.mark.gpu


def make(bits, V, signed, M, N, K, sparsity, seed, mask="poisson"):
    gen = synthetic.poisson_block_mask if mask == "poisson" else synthetic.uniform_block_mask
    csr = gen(M, N, V, sparsity, seed)
    lhs = synthetic.random_lanes(M * K, bits, signed, seed + 1)
    rhs = synthetic.random_lanes(N * K, bits, signed, seed + 2)
    return csr, lhs, rhs


def run(bits, V, signed, csr, lhs, rhs, K, layout=sddmm.OUT_REFERENCE, prefill=0x5A5A5A5A):
    dm = sddmm.DeviceMask.upload(csr, alignment=8)
    l = torch.from_numpy(formats.pack_lanes(lhs, bits).view(np.int32)).cuda()
    r = torch.from_numpy(formats.pack_lanes(rhs, bits).view(np.int32)).cuda()
    out = torch.full((dm.output_size(),), prefill, dtype=torch.int32, device="cuda")
    print(f"sddmm bits={bits} V={V} signed={signed} m_vec={csr.m_vec} N={csr.k} K={K} layout={layout}", flush=True)
    dm.sddmm(bits, l, r, K, out=out, sign=sddmm.SIGNED if signed else sddmm.UNSIGNED, layout=layout)
    torch.cuda.synchronize()
    return dm, out.cpu().numpy()


def expected(orc, dm, csr, lhs, rhs, K, V):
    exp = np.full(dm.output_size(), 0x5A5A5A5A, np.int32)
    # written region: real slots and the padding slots up to the next multiple of 8
    for i in range(csr.m_vec):
        s, e = dm.pairs[2 * i], dm.pairs[2 * i + 1]
        for j in range(s, s + ((e - s + 7) & ~7)):
            for v in range(V):
                exp[(j // 8) * 8 * V + 8 * v + j % 8] = 0
    got = orc.sddmm_plain(csr.m_vec, V, K, dm.pairs, dm.cols, lhs, rhs, dm.output_size(), 8)
    mask = np.zeros(dm.output_size(), bool)
    for i in range(csr.m_vec):
        s, e = dm.pairs[2 * i], dm.pairs[2 * i + 1]
        for j in range(s, e):
            for v in range(V):
                mask[(j // 8) * 8 * V + 8 * v + j % 8] = True
    exp[mask] = got[mask]
    return exp


# 192: a pass with an empty chunk; 640: a full pass then a partial one
@pytest.mark.parametrize("K", [64, 128, 192, 256, 512, 640, 1024])
@pytest.mark.parametrize("signed", [True, False])
@pytest.mark.parametrize("V", [2, 4, 8])
@pytest.mark.parametrize("bits", [8, 4, 16])
def test_sddmm_small_exact(orc, bits, V, signed, K):
    csr, lhs, rhs = make(bits, V, signed, M=24 * V, N=300, K=K, sparsity=0.8, seed=V * 7 + K + bits)
    dm, got = run(bits, V, signed, csr, lhs, rhs, K)
    assert np.array_equal(got, expected(orc, dm, csr, lhs, rhs, K, V))


@pytest.mark.parametrize("bits", [8, 4, 16])
def test_sddmm_vector_layout(orc, bits):
    V, K = 8, 128
    csr, lhs, rhs = make(bits, V, True, M=256, N=512, K=K, sparsity=0.9, seed=3)
    dm, ref = run(bits, V, True, csr, lhs, rhs, K, sddmm.OUT_REFERENCE)
    _, vec = run(bits, V, True, csr, lhs, rhs, K, sddmm.OUT_VECTOR)
    for i in range(csr.m_vec):
        s, e = dm.pairs[2 * i], dm.pairs[2 * i + 1]
        for j in range(s, e):
            for v in range(V):
                assert vec[j * V + v] == ref[(j // 8) * 8 * V + 8 * v + j % 8]


@pytest.mark.parametrize("bits", [8, 4, 16])
def test_sddmm_reference_generator_unsigned(orc, bits):
    # MakeDenseMatrix<int> operands, checked against the literal host loop
    V, M, N, K = 8, 512, 512, 128
    csr = synthetic.uniform_block_mask(M, N, V, 0.9, seed=1)
    pairs, an = orc.align_row_offsets(csr.row_offsets, 8)
    cols = orc.align_col_indices(csr.row_offsets, csr.col_indices, pairs, an)
    lhs_w = orc.make_dense_matrix(4, M * K * bits // 32)
    rhs_w = orc.make_dense_matrix(4, N * K * bits // 32)
    exp, _ = orc.sddmm_ref_check(lhs_w, rhs_w, M, K, N, bits, bits, V, pairs, cols, csr.m_vec, 8, an * V)
    dm = sddmm.DeviceMask.upload(csr, alignment=8)
    out = torch.zeros(dm.output_size(), dtype=torch.int32, device="cuda")
    dm.sddmm(bits, torch.from_numpy(lhs_w).cuda(), torch.from_numpy(rhs_w).cuda(), K, out=out, sign=sddmm.UNSIGNED)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy()[: an * V], exp)


@pytest.mark.parametrize("sparsity", [0.9, 0.95])
@pytest.mark.parametrize("K", [64, 128])
@pytest.mark.parametrize("bits", [8, 4, 16])
def test_sddmm_full_size_exact(orc, bits, K, sparsity):
    # Q 4096 x K, K-matrix 4096 x K, V=8 mask: the SDDMM workload
    V = 8
    csr, lhs, rhs = make(bits, V, True, M=4096, N=4096, K=K, sparsity=sparsity, seed=11, mask="uniform")
    dm, got = run(bits, V, True, csr, lhs, rhs, K)
    exp_plain = orc.sddmm_plain(csr.m_vec, V, K, dm.pairs, dm.cols, lhs, rhs, dm.output_size(), 8)
    # compare on real slots (every row here has a multiple of 8 real slots or padding zeros)
    real = np.zeros(dm.output_size(), bool)
    idx = np.concatenate([np.arange(dm.pairs[2 * i], dm.pairs[2 * i + 1]) for i in range(csr.m_vec)])
    for v in range(V):
        real[(idx // 8) * 8 * V + 8 * v + idx % 8] = True
    assert np.array_equal(got[real], exp_plain[real])


@pytest.mark.parametrize("signed", [True, False])
def test_sddmm_16b_extreme_values(orc, signed):
    # every lane at the range ends: the offset terms (K * OA * OB) wrap mod 2^32
    V, K = 8, 256
    csr, lhs, rhs = make(16, V, signed, M=64, N=96, K=K, sparsity=0.5, seed=21)
    lo, hi = (-(1 << 15), (1 << 15) - 1) if signed else (0, (1 << 16) - 1)
    lhs = np.where(lhs % 2 == 0, lo, hi).astype(np.int32)
    rhs = np.where(rhs % 3 == 0, lo, hi).astype(np.int32)
    dm, got = run(16, V, signed, csr, lhs, rhs, K)
    assert np.array_equal(got, expected(orc, dm, csr, lhs, rhs, K, V))


@pytest.mark.parametrize("bits", [8, 16])
def test_sddmm_published_k512_shape_exact(orc, bits):
    # the reference's published SDDMM configuration: 256 vector rows of V=8,
    # N = 2304, K = 512, 90% sparse (one pass of eight 64-deep chunks)
    V, K = 8, 512
    csr, lhs, rhs = make(bits, V, False, M=2048, N=2304, K=K, sparsity=0.9, seed=41)
    dm, got = run(bits, V, False, csr, lhs, rhs, K)
    assert np.array_equal(got, expected(orc, dm, csr, lhs, rhs, K, V))
