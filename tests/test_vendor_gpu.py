"""Vendor-library comparators on the GPU: rocSPARSE and hipBLASLt results on
the benchmark operands agree with this library's SpMM / SDDMM and with the
oracle.

Integer paths (rocSPARSE int8 CSR / BSR with int32 output, hipBLASLt int8
GEMM) must match bit for bit.  The fp16 paths are exact where every partial
sum is an integer below 2^24 (operands in the int4 range, fp32 accumulation);
hipBLASLt's fp16 output is compared within half an fp16 ulp of the exact
result.  A combination the installed library does not provide (status
"unsupported") skips with its message.
"""
import numpy as np
import pytest
import torch

from magicube_amd import sddmm, spmm, synthetic, vendor
from tests.spmm_cases import make_case

pytestmark = pytest.mark.gpu

UNSUPPORTED = 6


def plan_or_skip(make):
    try:
        return make()
    except vendor.VendorError as exc:
        if exc.status == UNSUPPORTED:
            pytest.skip(f"not provided by the installed library: {exc}")
        raise


def ours(case):
    dev = spmm.DeviceSpmm.upload(case.prep)
    rhs = torch.from_numpy(case.rhs_packed).cuda()
    out = torch.full((case.m_vec * case.V, case.N), 0x5A5A5A5A, dtype=torch.int32, device="cuda")
    print(f"ours L{case.lhs_bits}-R{case.rhs_bits} V={case.V} m_vec={case.m_vec} K={case.K} N={case.N}", flush=True)
    dev.matmul(rhs, case.N, out=out, sign=spmm.SIGNED if case.signed else spmm.UNSIGNED)
    torch.cuda.synchronize()
    return out


def operands(case):
    a8 = case.a.astype(np.int8)
    b8 = torch.from_numpy(case.b.astype(np.int8).reshape(case.K, case.N)).cuda()
    return a8, b8


SHAPES = [  # (V, M, K, N, sparsity)
    (8, 512, 512, 128, 0.9),
    (4, 384, 768, 256, 0.7),
    (2, 256, 1024, 64, 0.98),
    (8, 4096, 4096, 512, 0.9),
]


@pytest.mark.parametrize("alg", list(vendor.CSR_ALGS))
@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: f"V{s[0]}_M{s[1]}_K{s[2]}_N{s[3]}")
def test_rocsparse_csr_int8_matches(orc, shape, alg):
    V, M, K, N, s = shape
    case = make_case(8, 8, V, True, M=M, K=K, N=N, sparsity=s, seed=5 + V)
    got = ours(case)
    a8, b8 = operands(case)
    rp, ci, gather = vendor.scalar_csr(case.csr)
    rp_t, ci_t = torch.from_numpy(rp).cuda(), torch.from_numpy(ci).cuda()
    vals = torch.from_numpy(np.ascontiguousarray(a8[gather])).cuda()
    c = torch.full((M, N), 0x3C3C3C3C, dtype=torch.int32, device="cuda")
    print(f"rocsparse csr {alg} V={V} M={M} K={K} N={N}", flush=True)
    plan = plan_or_skip(lambda: vendor.csr_spmm(vendor.I8, alg, M, K, N, rp_t, ci_t, vals, b8, c))
    plan.run()
    torch.cuda.synchronize()
    assert torch.equal(c, got)
    if M * N <= 512 * 128:
        assert np.array_equal(c.cpu().numpy(), orc.spmm_plain(case.m_vec, V, N, case.csr.row_offsets,
                                                              case.csr.col_indices, case.a, case.b))
    plan.close()


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: f"V{s[0]}_M{s[1]}_K{s[2]}_N{s[3]}")
def test_rocsparse_bsr_int8_matches(shape):
    V, M, K, N, s = shape
    case = make_case(8, 8, V, True, M=M, K=K, N=N, sparsity=s, seed=9 + V)
    got = ours(case)
    a8, b8 = operands(case)
    brp, bci, bv = vendor.bsr_blocks(case.csr, a8)
    t = [torch.from_numpy(x).cuda() for x in (brp, bci, bv)]
    c = torch.full((M, N), 0x3C3C3C3C, dtype=torch.int32, device="cuda")
    print(f"rocsparse bsr V={V} M={M} K={K} N={N} nnzb={bci.size}", flush=True)
    plan = plan_or_skip(lambda: vendor.bsr_spmm(vendor.I8, case.m_vec, K // V, N, V, *t, b8, c))
    plan.run()
    torch.cuda.synchronize()
    assert torch.equal(c, got)


def test_rocsparse_bsr_uniform_blocks_exact():
    """The reference's Blocked-ELL baseline structure, checked against an
    int64 dense product on the host."""
    m_vec, K, N, V = 64, 512, 128, 8
    rp, ci = vendor.uniform_blocks(m_vec, K, V, 0.1, seed=3)
    vals = synthetic.random_lanes(ci.size * V * V, 8, True, 4).astype(np.int8)
    b = synthetic.random_lanes(K * N, 8, True, 5).astype(np.int8).reshape(K, N)
    dense = np.zeros((m_vec * V, K), np.int64)
    blocks = vals.reshape(-1, V, V)
    for i in range(m_vec):
        for t in range(rp[i], rp[i + 1]):
            dense[i * V:(i + 1) * V, ci[t] * V:(ci[t] + 1) * V] = blocks[t]
    exp = (dense @ b.astype(np.int64)).astype(np.int32)
    t = [torch.from_numpy(x).cuda() for x in (rp, ci, vals)]
    bt = torch.from_numpy(b).cuda()
    c = torch.zeros((m_vec * V, N), dtype=torch.int32, device="cuda")
    plan = plan_or_skip(lambda: vendor.bsr_spmm(vendor.I8, m_vec, K // V, N, V, *t, bt, c))
    plan.run()
    torch.cuda.synchronize()
    assert np.array_equal(c.cpu().numpy(), exp)


@pytest.mark.parametrize("b_transposed", [False, True], ids=["NN", "TN"])
@pytest.mark.parametrize("shape", [SHAPES[0], SHAPES[3]], ids=["small", "headline"])
def test_hipblaslt_int8_gemm_matches(shape, b_transposed):
    V, M, K, N, s = shape
    case = make_case(8, 8, V, True, M=M, K=K, N=N, sparsity=s, seed=13)
    got = ours(case)
    a8, b8 = operands(case)
    dense = torch.from_numpy(vendor.dense_lhs(case.csr, a8)).cuda()
    b = b8.t().contiguous() if b_transposed else b8
    c = torch.full((M, N), 0x3C3C3C3C, dtype=torch.int32, device="cuda")
    print(f"hipblaslt int8 {'TN' if b_transposed else 'NN'} M={M} K={K} N={N}", flush=True)
    plan = plan_or_skip(lambda: vendor.gemm(vendor.I8, M, K, N, dense, b, c, b_transposed))
    plan.run()
    torch.cuda.synchronize()
    assert torch.equal(c, got)


def small_int_case(V=8, M=1024, K=1024, N=256, s=0.9, seed=21):
    """Operands in [-8, 7]: every fp32 partial sum is an exact integer."""
    case = make_case(8, 8, V, True, M=M, K=K, N=N, sparsity=s, seed=seed)
    rng = np.random.default_rng(seed)
    case.a[:] = rng.integers(-8, 8, size=case.a.size)
    case.b[:] = rng.integers(-8, 8, size=case.b.size)
    return make_case(8, 8, V, True, M=M, K=K, N=N, sparsity=s, seed=seed, a=case.a, b=case.b)


@pytest.mark.parametrize("alg", list(vendor.CSR_ALGS))
def test_rocsparse_csr_fp16_exact_on_small_integers(alg):
    case = small_int_case()
    got = ours(case)
    a8, b8 = operands(case)
    rp, ci, gather = vendor.scalar_csr(case.csr)
    vals = torch.from_numpy(a8[gather].astype(np.float16)).cuda()
    c = torch.zeros((case.m_vec * case.V, case.N), dtype=torch.float32, device="cuda")
    plan = plan_or_skip(lambda: vendor.csr_spmm(vendor.F16, alg, case.m_vec * case.V, case.K, case.N,
                                                torch.from_numpy(rp).cuda(), torch.from_numpy(ci).cuda(), vals,
                                                b8.to(torch.float16), c))
    plan.run()
    torch.cuda.synchronize()
    assert torch.equal(c, got.to(torch.float32))


def test_hipblaslt_fp16_gemm_within_half_ulp():
    case = small_int_case()
    got = ours(case).to(torch.float64)
    a8, b8 = operands(case)
    M, K, N = case.m_vec * case.V, case.K, case.N
    dense = torch.from_numpy(vendor.dense_lhs(case.csr, a8).astype(np.float16)).cuda()
    c = torch.zeros((M, N), dtype=torch.float16, device="cuda")
    plan = plan_or_skip(lambda: vendor.gemm(vendor.F16, M, K, N, dense, b8.t().contiguous().to(torch.float16), c,
                                            True))
    plan.run()
    torch.cuda.synchronize()
    # fp16 spacing at |x| is 2^(floor(log2|x|) - 10); round-to-nearest is within half of it
    ulp = torch.exp2(torch.floor(torch.log2(got.abs().clamp(min=1.0))) - 10)
    assert bool(((c.to(torch.float64) - got).abs() <= 0.5 * ulp).all())


@pytest.mark.parametrize("V", [8, 4, 2])
def test_rocsparse_sddmm_fp16_matches(V):
    M = N = 512
    K = 64
    csr = synthetic.uniform_block_mask(M, N, V, 0.9, seed=31 + V)
    mask = sddmm.DeviceMask.upload(csr)
    rng = np.random.default_rng(V)
    la = rng.integers(-8, 8, size=M * K).astype(np.int8)
    lb = rng.integers(-8, 8, size=N * K).astype(np.int8)
    ours_out = mask.sddmm(8, torch.from_numpy(la.view(np.int32)).cuda(), torch.from_numpy(lb.view(np.int32)).cuda(),
                          K, layout=sddmm.OUT_VECTOR)
    rp, ci, _ = vendor.scalar_csr(csr)
    out = torch.ones(ci.size, dtype=torch.float32, device="cuda")
    plan = plan_or_skip(lambda: vendor.csr_sddmm(M, N, K, torch.from_numpy(rp).cuda(), torch.from_numpy(ci).cuda(),
                                                 torch.from_numpy(la.astype(np.float16).reshape(M, K)).cuda(),
                                                 torch.from_numpy(lb.astype(np.float16).reshape(N, K)).cuda(), out))
    print(f"rocsparse sddmm V={V} nnz={ci.size}", flush=True)
    plan.run()
    torch.cuda.synchronize()
    # this library writes [aligned slot][v]; the vendor writes scalar-CSR order
    ours_np = ours_out.cpu().numpy().reshape(-1, V)
    slots = np.concatenate([np.arange(mask.pairs[2 * i], mask.pairs[2 * i] + csr.row_offsets[i + 1] -
                                      csr.row_offsets[i]) for i in range(csr.m_vec)])
    vec_major = ours_np[slots].reshape(-1)
    scalar = out.cpu().numpy()[vendor.vector_to_scalar_index(csr)]
    assert np.array_equal(scalar.astype(np.int64), vec_major.astype(np.int64))
