"""B-stationary SpMM configs (8-bit B) on the GPU: bit-exact against the
oracle for both 8-bit-B pairs, every V and signedness, empty and
single-vector rows, every row order, K at the LDS limit, and the headline
shape by checksum; shapes that do not fit are rejected."""
import numpy as np
import pytest
import torch

from magicube_amd import _lib, formats, spmm, synthetic
from tests.spmm_cases import checksum_expected, checksum_got, make_case
from tests.test_spmm_gpu import oracle_result, run_case

pytestmark = pytest.mark.gpu


def stationary_configs():
    return [i for i, c in enumerate(spmm.tile_configs()) if c[4] == 1]


@pytest.mark.parametrize("signed", [True, False])
@pytest.mark.parametrize("V", [2, 4, 8])
@pytest.mark.parametrize("pair", [(8, 8), (16, 8)])
def test_stationary_exact(orc, pair, V, signed):
    for cfg in stationary_configs():
        case = make_case(*pair, V, signed, M=80 * V, K=512, N=256, sparsity=0.85, seed=3 * cfg + V)
        got = run_case(case, config=cfg)
        assert np.array_equal(got, oracle_result(orc, case)), f"config {cfg}"


@pytest.mark.parametrize("row_order", ["sorted", "identity", "null"])
def test_stationary_empty_and_single_rows(orc, row_order):
    V, K, N = 8, 700, 128
    case = make_case(8, 8, V, True, M=37 * V, K=K, N=N, sparsity=0.9, seed=41)
    counts = np.diff(case.csr.row_offsets).copy()
    counts[[0, 5, 36]] = 0            # empty rows, including the last one
    counts[[1, 7]] = 1                # single nonzero vector
    rng = np.random.default_rng(4)
    cols = [np.sort(rng.choice(K, size=c, replace=False)) for c in counts]
    ro = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    csr = formats.VectorCSR(37, K, V, ro, np.concatenate(cols).astype(np.int32))
    a = synthetic.random_lanes(csr.nnz_vec * V, 8, True, 42)
    case.csr, case.a = csr, a
    case.prep = formats.prepare_spmm(csr, formats.pack_lanes(a, 8), 8, 8)
    for cfg in stationary_configs():
        got = run_case(case, config=cfg, row_order=row_order)
        assert np.array_equal(got, oracle_result(orc, case)), f"config {cfg}"


def test_stationary_k_at_the_lds_limit_and_rejections(orc):
    cfgs = {i: c for i, c in enumerate(spmm.tile_configs()) if c[4] == 1}
    limit = {i: (159 * 1024) // (4 * c[0]) - 1 for i, c in cfgs.items()}    # K + 1 rows fit
    for cfg, kmax in limit.items():
        case = make_case(8, 8, 8, True, M=24 * 8, K=kmax, N=4 * cfgs[cfg][0] * 2, sparsity=0.97, seed=cfg)
        got = run_case(case, config=cfg)
        assert np.array_equal(got, oracle_result(orc, case)), f"config {cfg} K={kmax}"
        big = make_case(8, 8, 8, True, M=16, K=kmax + 1, N=4 * cfgs[cfg][0], sparsity=0.99, seed=cfg)
        with pytest.raises(_lib.MagicubeError) as e:
            run_case(big, config=cfg)
        assert e.value.status == _lib.MC_EINVAL_SHAPE
    odd_n = make_case(8, 8, 8, True, M=16, K=256, N=48, sparsity=0.9, seed=2)
    with pytest.raises(_lib.MagicubeError):
        run_case(odd_n, config=min(cfgs))


@pytest.mark.parametrize("signed", [True, False])
def test_stationary_headline_checksum(signed):
    case = make_case(8, 8, 8, signed, M=4096, K=4096, N=512, sparsity=0.9, seed=7)
    w = np.random.default_rng(1).integers(0, 1 << 10, size=case.N)
    for cfg in stationary_configs():
        if (4096 + 1) * 4 * spmm.tile_configs()[cfg][0] > 159 * 1024:
            continue
        got = run_case(case, config=cfg)
        assert np.array_equal(checksum_got(got, w), checksum_expected(case, w)), f"config {cfg}"
