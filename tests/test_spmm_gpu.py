"""SpMM parity on the GPU through the C ABI: bit-exact against the oracle.

Exact comparisons at sizes the oracle finishes in about a second (including
the 512x512 CPU-path workload and the M=K=4096, N=512 headline), plus
size-independent checksums C w == A (B w) (mod 2^32) over the whole
precision sweep (pairs x V x sparsity x N at M=K=4096).
"""
import sys

import numpy as np
import pytest
import torch

from magicube_amd import formats, spmm, synthetic
from tests.spmm_cases import checksum_expected, checksum_got, make_case

pytestmark = pytest.mark.gpu

PAIRS = [(8, 8), (16, 8), (8, 4), (4, 4), (12, 8), (16, 16), (12, 4), (16, 4)]
BASE_PAIRS = PAIRS[:4]


def pair_id(p):
    return f"L{p[0]}R{p[1]}" if isinstance(p, tuple) else None


def run_case(case, config=None, row_order="sorted", prefill=0x5A5A5A5A, stream=None):
    dev = spmm.DeviceSpmm.upload(case.prep)
    if row_order == "identity":
        dev.row_indices = torch.arange(case.m_vec, dtype=torch.int32, device="cuda")
    elif row_order == "null":
        dev.row_indices = None
    rhs = torch.from_numpy(case.rhs_packed).cuda()
    out = torch.full((case.m_vec * case.V, case.N), prefill, dtype=torch.int32, device="cuda")
    sign = spmm.SIGNED if case.signed else spmm.UNSIGNED
    print(f"case L{case.lhs_bits}-R{case.rhs_bits} V={case.V} signed={case.signed} m_vec={case.m_vec} "
          f"K={case.K} N={case.N} cfg={config}", flush=True)
    dev.matmul(rhs, case.N, out=out, sign=sign, config=config, stream=stream)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def oracle_result(orc, case):
    return orc.spmm_plain(case.m_vec, case.V, case.N, case.csr.row_offsets, case.csr.col_indices, case.a,
                          case.b)


@pytest.mark.parametrize("signed", [True, False])
@pytest.mark.parametrize("V", [2, 4, 8])
@pytest.mark.parametrize("pair", PAIRS, ids=pair_id)
def test_spmm_small_exact(orc, pair, V, signed):
    case = make_case(*pair, V, signed, M=96 * V // 2, K=384, N=256, sparsity=0.8, seed=11 * V + pair[0])
    got = run_case(case)
    assert np.array_equal(got, oracle_result(orc, case))


GATHER_CONFIGS = [0, 1, 2, 3, 4, 5, 6, 7, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23]


@pytest.mark.parametrize("config", GATHER_CONFIGS)
@pytest.mark.parametrize("pair", PAIRS, ids=pair_id)
def test_spmm_every_tile_config(orc, pair, config):
    assert spmm.tile_configs()[config][4] == 0
    if config in (2, 3, 6) and pair[1] == 4:
        pytest.skip("four dwords per lane are built for 8- and 16-bit B only")
    for V in (2, 8):
        case = make_case(*pair, V, True, M=64 * V, K=512, N=512, sparsity=0.85, seed=config * 7 + V)
        got = run_case(case, config=config)
        assert np.array_equal(got, oracle_result(orc, case)), f"V={V}"


@pytest.mark.parametrize("pair", PAIRS, ids=pair_id)
def test_spmm_ragged_n_and_empty_rows(orc, pair):
    # N not a multiple of the tile; rows with 0, 1 and many nonzeros
    cpd = 32 // pair[1]
    N = 13 * cpd * 3
    case = make_case(*pair, 4, True, M=40 * 4, K=300, N=N, sparsity=0.9, seed=5)
    ro = case.csr.row_offsets
    assert (np.diff(ro) >= 0).all()
    got = run_case(case)
    assert np.array_equal(got, oracle_result(orc, case))


def test_spmm_fully_overwrites_output_and_handles_empty_matrix(orc):
    csr = formats.VectorCSR(8, 64, 8, np.zeros(9, np.int32), np.zeros(0, np.int32))
    prep = formats.prepare_spmm(csr, np.zeros(0, np.uint8), 8, 8)
    dev = spmm.DeviceSpmm.upload(prep)
    rhs = torch.ones(64 * 64 // 4, dtype=torch.int32, device="cuda")
    out = torch.full((64, 64), 12345, dtype=torch.int32, device="cuda")
    dev.matmul(rhs, 64, out=out)
    torch.cuda.synchronize()
    assert (out == 0).all()


@pytest.mark.parametrize("order", ["sorted", "identity", "null"])
def test_spmm_row_order_does_not_change_results(orc, order):
    case = make_case(8, 8, 8, True, M=512, K=512, N=128, sparsity=0.9, seed=3)
    assert np.array_equal(run_case(case, row_order=order), oracle_result(orc, case))


def test_spmm_side_stream(orc):
    case = make_case(8, 4, 8, True, M=256, K=512, N=256, sparsity=0.9, seed=4)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        got = run_case(case, stream=s)
    assert np.array_equal(got, oracle_result(orc, case))


@pytest.mark.parametrize("signed", [True, False])
def test_cpu_path_workload_512(orc, signed):
    # 512x512 A, V=8, 90% sparse with 51 vectors per block, B 512x128 int8
    case = make_case(8, 8, 8, signed, M=512, K=512, N=128, sparsity=0.9, seed=2024, mask="uniform")
    assert np.all(np.diff(case.csr.row_offsets) == 51)
    assert np.array_equal(run_case(case), oracle_result(orc, case))


def test_reference_generator_data_unsigned(orc):
    # Values from the reference's own generator (MakeDenseMatrix<long long> /
    # <int>) checked against the literal compute_ref_integers restatement.
    M, K, N, V = 512, 512, 128, 8
    csr = synthetic.uniform_block_mask(M, K, V, 0.9, seed=1)
    nnz = csr.nnz_vec
    vals = orc.make_dense_matrix(8, nnz)            # one long long per V=8 vector
    B = orc.make_dense_matrix(4, K * N // 4)
    C_ref, _ = orc.spmm_ref_check(vals, 8, B, M, K, N, 8, 8, 8, V, csr.row_offsets, csr.col_indices,
                                  csr.m_vec, 1)
    prep = formats.prepare_spmm(csr, vals.view(np.uint8), 8, 8)
    dev = spmm.DeviceSpmm.upload(prep)
    out = torch.empty((M, N), dtype=torch.int32, device="cuda")
    dev.matmul(torch.from_numpy(B).cuda(), N, out=out, sign=spmm.UNSIGNED)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    assert np.array_equal(got, C_ref)
    assert (got > 0).sum() == M * N   # the published counter = M*N invariant


@pytest.mark.parametrize("signed", [True, False])
def test_headline_exact(orc, signed):
    # M=K=4096, N=512, V=8, 90% sparse (Poisson per block), L8-R8
    case = make_case(8, 8, 8, signed, M=4096, K=4096, N=512, sparsity=0.9, seed=7)
    assert np.array_equal(run_case(case), oracle_result(orc, case))


SWEEP = ([(pair, V, s, N) for pair in BASE_PAIRS for V in (2, 4, 8) for s in (0.7, 0.9, 0.98) for N in (128, 256, 512)]
         + [(pair, V, 0.9, 256) for pair in PAIRS[4:] for V in (2, 4, 8)])


@pytest.mark.parametrize("pair,V,s,N", SWEEP, ids=lambda x: pair_id(x) if isinstance(x, tuple) else str(x))
def test_sweep_checksums(pair, V, s, N):
    case = make_case(*pair, V, True, M=4096, K=4096, N=N, sparsity=s, seed=int(s * 100) + V + N)
    got = run_case(case)
    rng = np.random.default_rng(0)
    for _ in range(2):
        w = rng.integers(0, 1 << 10, size=N)
        assert np.array_equal(checksum_got(got, w), checksum_expected(case, w))


@pytest.mark.parametrize("signed", [True, False])
@pytest.mark.parametrize("pair", PAIRS[4:], ids=pair_id)
def test_spmm_extreme_values(orc, pair, signed):
    # lanes at the ends of their ranges: digit offsets and weights at their limits
    case = make_case(*pair, 8, signed, M=256, K=512, N=128, sparsity=0.7, seed=99)
    la, lb = pair
    def ends(x, bits):
        lo, hi = (-(1 << (bits - 1)), (1 << (bits - 1)) - 1) if signed else (0, (1 << bits) - 1)
        return np.where(x % 2 == 0, lo, hi).astype(np.int32)
    case.a = ends(case.a, la)
    case.b = ends(case.b, lb)
    case.prep = formats.prepare_spmm(case.csr, formats.pack_lanes(case.a, la), la, lb)
    rhs = formats.pack_lanes(case.b, lb)
    case.rhs_packed = np.concatenate([rhs, np.zeros((-rhs.size) % 4, np.uint8)]).view(np.int32)
    assert np.array_equal(run_case(case), oracle_result(orc, case))


def _default_cfg(lhs, rhs, V, n, m_vec):
    """The library's default tile rule, restated (spmm_api.hip default_config)."""
    tc = 16 * (32 // rhs)
    tiles = -(-n // tc)
    if tiles >= 8 and 4096 <= m_vec * (-(-tiles // 8)) * 8 < 8192:
        return 13
    tpb = 4 if tiles >= 4 else (2 if tiles >= 2 else 1)
    if m_vec * (-(-tiles // tpb)) * tpb < 2048:
        return 20 if tpb == 4 else 22
    return {4: 17, 2: 16, 1: 15}[tpb]


def _default_cases():
    """Seeded shapes, eight per default tile (13, 15, 16, 17, 20, 22): ragged
    N (any multiple of 32/rhs_bits), row counts on both sides of the 2048-,
    4096- and 8192-wave splits, every pair and V."""
    rng = np.random.default_rng(2024)
    cases = []
    targets = (13, 15, 16, 17, 20, 22, 17)
    for i in range(8 * len(targets)):
        target = targets[i % len(targets)]
        pair = PAIRS[(i // len(targets)) % len(PAIRS)]
        V = (2, 4, 8)[i % 3]
        cpd = 32 // pair[1]
        r = int(rng.integers(1, 17) if target == 15 else
                rng.integers(17, 49) if target in (16, 22) else
                rng.integers(113, 161) if target == 13 or i % len(targets) == 6 else rng.integers(49, 97))
        n = cpd * r
        tiles = -(-r // 16)
        tpb = 8 if target == 13 else 4 if tiles >= 4 else (2 if tiles >= 2 else 1)
        per_row = -(-tiles // tpb) * tpb
        if target == 13:
            m_vec = int(rng.integers(-(-4096 // per_row), 8192 // per_row))
        elif i % len(targets) == 6:
            # eight or more tiles past the 8192-wave split: back to four waves per row
            m_vec = -(-8192 // (-(-tiles // 8) * 8)) + int(rng.integers(0, 100))
        elif target in (20, 22):
            m_vec = int(rng.integers(1, 2048 // per_row))
        else:
            m_vec = -(-2048 // per_row) + int(rng.integers(0, 200))
        K = int(rng.integers(40, 600))
        s = float(rng.choice([0.5, 0.8, 0.95]))
        cases.append((pair, V, m_vec, K, n, s, i % 2 == 0, 1000 + i))
    return cases


def test_default_rule_cases_cover_every_default_tile():
    got = {_default_cfg(*c[0], c[1], c[4], c[2]) for c in _default_cases()}
    assert got == {13, 15, 16, 17, 20, 22}, got


@pytest.mark.parametrize("case", _default_cases(), ids=lambda c: f"L{c[0][0]}R{c[0][1]}-V{c[1]}-m{c[2]}-K{c[3]}-N{c[4]}-{'s' if c[6] else 'u'}")
def test_spmm_default_tiles_exact(orc, case):
    (la, lb), V, m_vec, K, N, s, signed, seed = case
    c = make_case(la, lb, V, signed, M=m_vec * V, K=K, N=N, sparsity=s, seed=seed)
    got = run_case(c)
    assert np.array_equal(got, oracle_result(orc, c)), f"default tile {_default_cfg(la, lb, V, N, m_vec)}"
