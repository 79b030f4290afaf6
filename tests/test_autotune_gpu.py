"""Autotuner on the GPU: every timed config agrees bitwise with the default,
the winner is installed on the tuned operand and cached."""
import numpy as np
import pytest
import torch

from magicube_amd import attention, autotune, formats, spmm, synthetic

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("bits", [(8, 8), (8, 4)])
def test_tune_spmm_installs_and_caches(bits, tmp_path):
    M = K = 1024
    N, V = 256, 8
    csr = synthetic.poisson_block_mask(M, K, V, 0.9, seed=4)
    a = synthetic.random_lanes(csr.nnz_vec * V, bits[0], True, 5)
    b = synthetic.random_lanes(K * N, bits[1], True, 6)
    dev = spmm.DeviceSpmm.upload(formats.prepare_spmm(csr, formats.pack_lanes(a, bits[0]), *bits))
    rhs = torch.from_numpy(formats.pack_lanes(b, bits[1]).view(np.int32)).cuda()
    cache = str(tmp_path / "tune.json")
    print("tune", bits, flush=True)
    res = autotune.tune_spmm(dev, rhs, N, launches=5, reps=3, cache_file=cache)
    assert not res.cached and res.config in res.timings and len(res.timings) >= 5
    assert dev.configs[N] == res.config
    fresh = spmm.DeviceSpmm.upload(dev.prep)
    again = autotune.tune_spmm(fresh, rhs, N, cache_file=cache)
    assert again.cached and again.config == res.config and fresh.configs[N] == res.config
    # the tuned operand runs the tuned config; results unchanged, and an
    # untuned operand still runs the default
    out = dev.matmul(rhs, N)
    ref = dev.matmul(rhs, N, config=res.config)
    other = spmm.DeviceSpmm.upload(dev.prep)
    assert other.configs == {}
    base = other.matmul(rhs, N)
    torch.cuda.synchronize()
    assert torch.equal(out, ref) and torch.equal(out, base)


def test_tune_attention_grid():
    mats = []
    for n, L in enumerate((512, 1024)):
        q, k, v = synthetic.attention_inputs(L, n)
        mats.append(attention.AttentionMatrix(*(torch.from_numpy(x).cuda() for x in (q, k, v)),
                                              synthetic.attention_mask(L, 8, 0.9, n)))
    b = attention.AttentionBatch(mats)
    best, timings = autotune.tune_attention(b, candidates=(0, 1024, 4096))
    assert best in timings and all(t > 0 for t in timings.values())
