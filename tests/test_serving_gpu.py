"""Continuous-batching server on the GPU: outputs identical alone and in the
mix (and to the one-op-per-launch pipeline), admission within the HBM
budget, oversize requests rejected."""
import math

import numpy as np
import pytest
import torch

from magicube_amd import attention, serving

pytestmark = pytest.mark.gpu


def small_trace(n=6, seed=3, rate=math.inf):
    return serving.make_requests(n, rate, seed, heads=2, samples=(1, 2), seq=(256, 1024), sparsity=(0.8, 0.95))


def _bits(outs):
    return [o.view(torch.int32).cpu().numpy() for o in outs]


def test_alone_equals_mixed_and_plain():
    reqs = small_trace()
    budget = 3 * max(r.ws_bytes for r in reqs)
    print("trace:", [(r.rid, r.qk_bits, len(r.matrices), r.ws_bytes) for r in reqs], flush=True)
    stats = serving.Server(budget, "batched", keep_outputs=True).run(reqs)
    torch.cuda.synchronize()
    assert stats["peak_working_set_bytes"] <= budget
    served = {r.rid: _bits(r.result()) for r in reqs}
    # the mix: one fused launch over every request of a width
    mixed = {}
    for bits in (8, 4):
        group = [r for r in reqs if r.qk_bits == bits]
        blocks = [attention.ResidentBlock(r.matrices, r.scales, bits, 8, 127.0) for r in group]
        attention.BlockBatch(blocks).forward()
        torch.cuda.synchronize()
        for r, blk in zip(group, blocks):
            mixed[r.rid] = _bits(blk.outputs())
    for r in reqs:
        serving.reset([r])
        serving.Server(budget, "batched", keep_outputs=True).run([r])
        torch.cuda.synchronize()
        alone = _bits(r.result())
        for a, b, c in zip(alone, mixed[r.rid], served[r.rid]):
            assert np.array_equal(a, b) and np.array_equal(a, c), f"request {r.rid} differs alone vs mixed"
    serving.reset(reqs)
    serving.Server(budget, "plain", keep_outputs=True).run(reqs)
    torch.cuda.synchronize()
    for r in reqs:
        for a, b in zip(_bits(r.result()), mixed[r.rid]):
            assert np.array_equal(a, b), f"request {r.rid} differs fused vs plain"


def test_budget_and_poisson_trace():
    reqs = small_trace(n=10, seed=5, rate=2000.0)
    budget = max(r.ws_bytes for r in reqs)          # at most one large request at a time
    stats = serving.Server(budget, "batched").run(reqs)
    assert stats["peak_working_set_bytes"] <= budget
    assert all(r.done >= r.launched >= r.admitted >= r.arrival for r in reqs)
    assert stats["p99_ms"] >= stats["p50_ms"] > 0
    with pytest.raises(ValueError):
        serving.Server(budget // 2, "batched").run(reqs)
