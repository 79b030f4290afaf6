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


def _same(a, b):
    return len(a) == len(b) and all(torch.equal(x.view(torch.int32), y.view(torch.int32)) for x, y in zip(a, b))


def test_stated_shape_alone_equals_mixed_and_plain():
    """The serving trace at its stated shape: 12 heads x 1-4 samples,
    L ~ U[256, 8192], sparsity ~ U[0.8, 0.98], three int8 and three int4
    requests.  Every request's outputs are bitwise the same served in the
    mix, computed alone, and from the one-op-per-launch pipeline; the peak
    working set stays within a budget of two of the largest requests."""
    reqs = serving.make_requests(3, math.inf, 21, qk_bits=(8,)) + serving.make_requests(3, math.inf, 22, qk_bits=(4,))
    for n, r in enumerate(reqs):
        r.rid = n
    print("trace:", [(r.rid, r.qk_bits, len(r.matrices), max(m.L for m in r.matrices)) for r in reqs], flush=True)
    assert max(m.L for r in reqs for m in r.matrices) > 4096
    budget = 2 * max(r.ws_bytes for r in reqs)
    stats = serving.Server(budget, "batched", keep_outputs=True).run(reqs)
    torch.cuda.synchronize()
    assert stats["peak_working_set_bytes"] <= budget
    served = {r.rid: [o.clone() for o in r.result()] for r in reqs}
    for r in reqs:
        serving.reset([r])
        serving.Server(budget, "fused", keep_outputs=True).run([r])
        torch.cuda.synchronize()
        assert _same(r.result(), served[r.rid]), f"request {r.rid} differs alone vs in the mix"
    serving.reset(reqs)
    serving.Server(budget, "plain", keep_outputs=True).run(reqs)
    torch.cuda.synchronize()
    for r in reqs:
        assert _same(r.result(), served[r.rid]), f"request {r.rid} differs batched vs plain"
    serving.reset(reqs)
