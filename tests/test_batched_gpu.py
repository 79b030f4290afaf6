"""Batched SpMM and SDDMM launches (mc_spmm_batched, mc_sddmm_batched) on the
GPU: every matrix of a batch bit-exact against the oracle, at byte strides
that leave gaps between the matrices (filled with a pattern the launch must
neither read as data nor overwrite)."""
import ctypes

import numpy as np
import pytest
import torch

from magicube_amd import _lib, formats, sddmm, spmm, synthetic
from tests.spmm_cases import make_case

pytestmark = pytest.mark.gpu

PAIRS = [(8, 8), (16, 8), (8, 4), (4, 4), (12, 8), (16, 16), (12, 4), (16, 4)]
GAP = 0x77   # byte pattern between matrices
PREFILL = 0x5A5A5A5A


def _p(t, byte_offset=0):
    return ctypes.c_void_p(t.data_ptr() + byte_offset)


def _stack(arrays, stride):
    """uint8 arrays laid out `stride` bytes apart, gaps filled with GAP."""
    buf = np.full(stride * len(arrays), GAP, np.uint8)
    for b, a in enumerate(arrays):
        a = np.ascontiguousarray(a).view(np.uint8)
        assert a.size <= stride
        buf[b * stride:b * stride + a.size] = a
    return torch.from_numpy(buf).cuda()


def _spmm_batch(orc, pair, V, signed, batch, config, seed):
    lhs_bits, rhs_bits = pair
    base = make_case(lhs_bits, rhs_bits, V, signed, M=48 * V, K=384, N=256, sparsity=0.8, seed=seed)
    cases = [base] + [make_case(lhs_bits, rhs_bits, V, signed, M=48 * V, K=384, N=256, sparsity=0.8, seed=seed,
                                a=synthetic.random_lanes(base.csr.nnz_vec * V, lhs_bits, signed, seed + 10 * b),
                                b=synthetic.random_lanes(384 * 256, rhs_bits, signed, seed + 10 * b + 1))
                      for b in range(1, batch)]
    for c in cases[1:]:   # one mask for the whole batch
        assert np.array_equal(c.prep.col_indices, base.prep.col_indices)
    dev = spmm.DeviceSpmm.upload(base.prep)
    v_stride = -(-base.prep.values.size // 16) * 16 + 48
    r_stride = base.rhs_packed.size * 4 + 64
    values = _stack([c.prep.values for c in cases], v_stride)
    rhs = _stack([c.rhs_packed for c in cases], r_stride)
    m, n = base.m_vec * V, base.N
    o_stride = m * n + 40
    out = torch.full((batch * o_stride,), PREFILL, dtype=torch.int32, device="cuda")
    print(f"batched spmm L{lhs_bits}R{rhs_bits} V={V} signed={signed} batch={batch} cfg={config} "
          f"m_vec={base.m_vec}", flush=True)
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.check(_lib.lib().mc_spmm_batched(lhs_bits, rhs_bits, spmm.SIGNED if signed else spmm.UNSIGNED, config,
                                          batch, base.m_vec, V, n, base.K, _p(dev.row_indices),
                                          _p(dev.row_offsets), _p(dev.col_indices), _p(values), v_stride, _p(rhs),
                                          r_stride, _p(out), o_stride, st), "mc_spmm_batched")
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    for b, c in enumerate(cases):
        exp = orc.spmm_plain(c.m_vec, V, n, c.csr.row_offsets, c.csr.col_indices, c.a, c.b)
        assert np.array_equal(got[b * o_stride:b * o_stride + m * n].reshape(m, n), exp), f"matrix {b}"
        assert np.all(got[b * o_stride + m * n:(b + 1) * o_stride] == PREFILL), f"gap after matrix {b}"


@pytest.mark.parametrize("signed", [True, False])
@pytest.mark.parametrize("V", [2, 8])
@pytest.mark.parametrize("pair", PAIRS, ids=lambda p: f"L{p[0]}R{p[1]}")
def test_spmm_batched_exact(orc, pair, V, signed):
    _spmm_batch(orc, pair, V, signed, batch=3, config=-1, seed=40 + pair[0] + pair[1] + V)


GATHER_CONFIGS = [c for c, t in enumerate(spmm.tile_configs()) if t[4] == 0]


@pytest.mark.parametrize("config", GATHER_CONFIGS)
@pytest.mark.parametrize("pair", [(8, 8), (8, 4)], ids=lambda p: f"L{p[0]}R{p[1]}")
def test_spmm_batched_every_gather_config(orc, pair, config):
    if spmm.tile_configs()[config][0] == 4 and pair[1] == 4:
        pytest.skip("four dwords per lane are built for 8- and 16-bit B only")
    _spmm_batch(orc, pair, 8, True, batch=2, config=config, seed=90 + config)


def _sddmm_inputs(bits, V, signed, seed, batch, M=64, N=300, K=128):
    csr = synthetic.poisson_block_mask(M * V // 8, N, V, 0.8, seed)
    lhs = [synthetic.random_lanes(csr.m_vec * V * K, bits, signed, seed + 10 * b + 1) for b in range(batch)]
    rhs = [synthetic.random_lanes(N * K, bits, signed, seed + 10 * b + 2) for b in range(batch)]
    return csr, lhs, rhs, K


# K = 640: a full pass of chunks, then a partial one accumulating into the output
@pytest.mark.parametrize("K", [128, 640])
@pytest.mark.parametrize("layout", [sddmm.OUT_REFERENCE, sddmm.OUT_VECTOR], ids=["ref", "vec"])
@pytest.mark.parametrize("signed", [True, False])
@pytest.mark.parametrize("V", [2, 8])
@pytest.mark.parametrize("bits", [8, 4, 16])
def test_sddmm_batched_exact(orc, bits, V, signed, layout, K):
    batch = 3
    csr, lhs, rhs, K = _sddmm_inputs(bits, V, signed, seed=7 * bits + V, batch=batch, K=K)
    dm = sddmm.DeviceMask.upload(csr, alignment=8)
    size = dm.output_size()
    l_stride = csr.m_vec * V * K * bits // 8 + 32
    r_stride = csr.k * K * bits // 8 + 16
    # an odd stride: matrices 1 and 2 start off 16-byte boundaries, so the
    # vector layout's wide stores give way to per-lane ones there
    o_stride = size + 25
    lt = _stack([formats.pack_lanes(x, bits) for x in lhs], l_stride)
    rt = _stack([formats.pack_lanes(x, bits) for x in rhs], r_stride)
    out = torch.full((batch * o_stride,), PREFILL, dtype=torch.int32, device="cuda")
    sign = sddmm.SIGNED if signed else sddmm.UNSIGNED
    print(f"batched sddmm bits={bits} V={V} signed={signed} layout={layout} K={K} m_vec={csr.m_vec}", flush=True)
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.check(_lib.lib().mc_sddmm_batched(bits, sign, layout, batch, csr.m_vec, K, csr.k, _p(dm.row_indices),
                                           _p(dm.row_offsets), _p(dm.col_indices), _p(lt), l_stride, _p(rt),
                                           r_stride, _p(out), o_stride, V, st), "mc_sddmm_batched")
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    for b in range(batch):
        exact = orc.sddmm_plain(csr.m_vec, V, K, dm.pairs, dm.cols, lhs[b], rhs[b], size, 8)
        mat = got[b * o_stride:b * o_stride + size]
        for i in range(csr.m_vec):
            s, e = dm.pairs[2 * i], dm.pairs[2 * i + 1]
            for j in range(s, s + ((e - s + 7) & ~7)):
                for v in range(V):
                    ref_at = (j // 8) * 8 * V + 8 * v + j % 8
                    at = ref_at if layout == sddmm.OUT_REFERENCE else j * V + v
                    assert mat[at] == (exact[ref_at] if j < e else 0), f"matrix {b} row {i} slot {j} lane {v}"
        assert np.all(got[b * o_stride + size:(b + 1) * o_stride] == PREFILL), f"gap after matrix {b}"
