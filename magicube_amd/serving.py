"""Continuous-batching server for quantized sparse-attention requests.

A request is one attention layer for 1-4 samples x H heads at its own
sequence length, with its own vector-sparse masks and static quantization
scales (the reference layer uses fixed scales, spattention.py:51-53).  Its
fp16 Q/K/V and masks are resident in HBM when it arrives, as the output of
the layers before it would be.  The server

* admits waiting requests first come, first served while the sum of their
  working sets (quantized Q/K/V and the fp32 outputs) fits an HBM budget,
  every request that fits in one pass; the batched policy builds one
  resident block (one arena, one vectorized descriptor build) per Q/K width
  for the whole admitted group, so the host work per request shrinks as
  groups grow;
* turns every admitted request that has not started into ONE step: a batched
  quantize launch per operand width and one persistent fused attention
  launch per Q/K width over the union of their (matrix, vector row) items;
* keeps up to ``max_inflight`` steps queued on alternating streams, so the
  host assembles the next step while the GPU runs the current one;
* retires the requests of a step when its event has completed, returns their
  budget and hands back the outputs.

Every item is computed by one wave from its own matrix's data, so a
request's outputs are bitwise the same alone or in any mix.

Two per-request policies run on the same loop for comparison:
``"fused"`` (one fused step per request) and ``"plain"`` (the
one-op-per-launch pipeline per matrix, as the reference layer launches it).
"""
from __future__ import annotations

import math
import statistics
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np
import torch

from . import attention, synthetic

POLICIES = ("batched", "fused", "plain")


@dataclass
class Request:
    rid: int
    arrival: float                      # seconds after the trace starts
    qk_bits: int
    matrices: list                      # [attention.AttentionMatrix] (samples x heads)
    scales: list                        # per matrix (scale_q, scale_k, scale_v)
    ws_bytes: int = 0                   # working set charged against the budget
    admitted: float = math.nan
    launched: float = math.nan
    done: float = math.nan
    outputs: list | None = None         # fp32 [L, 64] per matrix (plain policy)
    block: object = None                # attention.ResidentBlock holding the outputs (fused policies)
    block_range: tuple | None = None    # this request's matrices in a block shared with others

    def result(self) -> list:
        """fp32 [L, 64] output of every matrix (valid once the request is done)."""
        if self.block is None:
            return self.outputs
        outs = self.block.outputs()
        return outs if self.block_range is None else outs[self.block_range[0]:self.block_range[1]]

    @property
    def latency(self) -> float:
        return self.done - self.arrival


@dataclass
class _Step:
    reqs: list
    event: torch.cuda.Event
    keep: list = field(default_factory=list)     # batches and buffers alive until the step ends


def make_requests(n: int, rate: float, seed: int, heads: int = 12, samples=(1, 4), seq=(256, 8192),
                  sparsity=(0.8, 0.98), qk_bits=(8, 4), V: int = 8, device="cuda") -> list[Request]:
    """Synthetic request trace: Poisson arrivals at ``rate`` requests/s (all
    at t=0 when rate is inf), 1-4 samples, L ~ U[256, 8192] in multiples of
    64, sparsity ~ U[0.8, 0.98], int8 or int4 Q/K; inputs N(0, 1) fp16 and
    masks generated on the device; per-tensor scales from
    :func:`attention.qkv_scales`."""
    rng = np.random.default_rng(seed)
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    gaps = rng.exponential(1.0 / rate, size=n) if math.isfinite(rate) else np.zeros(n)
    arrivals = np.cumsum(gaps) - gaps[0]
    reqs = []
    for rid in range(n):
        ns = int(rng.integers(samples[0], samples[1] + 1))
        s = float(rng.uniform(*sparsity))
        bits = int(rng.choice(qk_bits))
        mats = []
        for _ in range(ns):
            L = int(rng.integers(seq[0] // 64, seq[1] // 64 + 1)) * 64
            for _ in range(heads):
                q, k, v = (torch.randn((L, attention.HEAD_DIM), generator=gen, device=device,
                                       dtype=torch.float32).half() for _ in range(3))
                mats.append(attention.AttentionMatrix(q, k, v, synthetic.attention_mask_device(L, V, s, gen, device)))
        scales = attention.qkv_scales(mats, bits)
        ws = sum(attention.working_set_bytes(m.L, bits) for m in mats)
        reqs.append(Request(rid, float(arrivals[rid]), bits, mats, scales, ws))
    return reqs


class Server:
    """Admission, step assembly and retirement (see the module docstring)."""

    def __init__(self, hbm_budget: int, policy: str = "batched", V: int = 8, p_mode: int = attention.P_ROW_SCALED,
                 p_scale: float = 127.0, max_inflight: int = 2, n_streams: int = 2,
                 grid_waves: int = attention.DEFAULT_GRID_WAVES,
                 keep_outputs: bool = False, max_step_items: int | None = None, device="cuda"):
        if policy not in POLICIES:
            raise ValueError(f"policy must be one of {POLICIES}")
        self.budget, self.policy, self.V = int(hbm_budget), policy, V
        self.p_mode, self.p_scale = p_mode, float(p_scale)
        self.max_inflight, self.grid_waves, self.keep_outputs = max_inflight, grid_waves, keep_outputs
        self.device = torch.device(device)
        self.max_step_items = max_step_items
        self.streams = [torch.cuda.Stream(device=self.device) for _ in range(n_streams)]
        self.prep_stream = torch.cuda.Stream(device=self.device)
        self.prep_event = torch.cuda.Event()
        self._next_stream = 0
        self.steps = 0
        self.peak_used = 0

    # -- admission and one step -------------------------------------------
    def _admit(self, group: list[Request], now: float) -> None:
        """Admits a group of requests whose working sets are reserved; for the
        fused policies also allocates their arenas and builds their
        descriptors (on the preparation stream, which every step waits for):
        one block per Q/K width for the whole group under the batched policy
        (unless steps are capped by ``max_step_items``, which may split a
        group), one block per request otherwise."""
        for r in group:
            r.admitted = now
        if self.policy == "plain":
            return
        with torch.cuda.stream(self.prep_stream):
            if self.policy == "batched" and self.max_step_items is None:
                by_width: dict[int, list] = {}
                for r in group:
                    by_width.setdefault(r.qk_bits, []).append(r)
                for bits, rs in by_width.items():
                    blk = attention.ResidentBlock([m for r in rs for m in r.matrices], [s for r in rs for s in r.scales],
                                                  bits, self.V, self.p_scale, self.device)
                    i = 0
                    for r in rs:
                        r.block, r.block_range = blk, (i, i + len(r.matrices))
                        i += len(r.matrices)
            else:
                for r in group:
                    r.block = attention.ResidentBlock(r.matrices, r.scales, r.qk_bits, self.V, self.p_scale,
                                                      self.device)
                    r.block_range = None
        self.prep_event.record(self.prep_stream)

    def _launch(self, reqs: list[Request], now: float) -> _Step:
        stream = self.streams[self._next_stream]
        self._next_stream = (self._next_stream + 1) % len(self.streams)
        keep = []
        with torch.cuda.stream(stream):
            for r in reqs:
                r.launched = now
            if self.policy == "plain":
                groups: dict[int, list] = {}
                for r in reqs:
                    devs = [attention.make_resident(m, sc, r.qk_bits, self.V, self.p_scale, self.device)
                            for m, sc in zip(r.matrices, r.scales)]
                    r.outputs = [d.o for d in devs]
                    groups.setdefault(r.qk_bits, []).extend(devs)
                for devs in groups.values():
                    batch = attention.AttentionBatch.from_resident(devs, V=self.V, p_mode=self.p_mode,
                                                                   p_scale=self.p_scale, device=self.device)
                    batch.forward_plain(stream)
                    keep.append(batch)
            else:
                stream.wait_event(self.prep_event)
                blocks: dict[int, list] = {}
                seen = set()
                for r in reqs:                      # requests of one group share a block
                    if id(r.block) not in seen:
                        seen.add(id(r.block))
                        blocks.setdefault(r.qk_bits, []).append(r.block)
                for blks in blocks.values():
                    batch = attention.BlockBatch(blks, p_mode=self.p_mode, p_scale=self.p_scale, device=self.device)
                    batch.forward(stream, self.grid_waves)
                    keep.append(batch)
            ev = torch.cuda.Event()
            ev.record(stream)
        self.steps += 1
        return _Step(reqs, ev, keep)

    def _take(self, ready: list[Request]) -> tuple[list[Request], list[Request]]:
        """The requests of the next step: one for the per-request policies;
        for the batched policy every ready request, up to ``max_step_items``
        work items (always at least one request)."""
        if self.policy != "batched":
            return ready[:1], ready[1:]
        if self.max_step_items is None:
            return ready, []
        n, items = 0, 0
        while n < len(ready):
            it = ready[n].block.n_items
            if n and items + it > self.max_step_items:
                break
            items += it
            n += 1
        return ready[:n], ready[n:]

    # -- the loop -------------------------------------------------------------
    def run(self, requests: list[Request]) -> dict:
        for r in requests:
            if r.ws_bytes > self.budget:
                raise ValueError(f"request {r.rid} needs {r.ws_bytes} bytes, more than the budget {self.budget}")
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        waiting: deque[Request] = deque()
        ready: list[Request] = []
        inflight: list[_Step] = []
        used = 0
        torch.cuda.synchronize(self.device)
        t0 = time.perf_counter()
        while pending or waiting or ready or inflight:
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())
            for st in [s for s in inflight if s.event.query()]:
                t = time.perf_counter() - t0
                inflight.remove(st)
                for r in st.reqs:
                    r.done = t
                    used -= r.ws_bytes
                    if not self.keep_outputs:
                        r.outputs = r.block = None
            if ready and len(inflight) < self.max_inflight:
                take, ready = self._take(ready)
                inflight.append(self._launch(take, now))
            # admit every waiting request that fits, in arrival order
            group = []
            while waiting and used + waiting[0].ws_bytes <= self.budget:
                r = waiting.popleft()
                used += r.ws_bytes
                group.append(r)
            if group:
                self._admit(group, now)
                ready.extend(group)
                self.peak_used = max(self.peak_used, used)
        span = max(r.done for r in requests) - min(r.arrival for r in requests) if requests else 0.0
        lat = sorted(r.latency for r in requests)
        return {
            "policy": self.policy,
            "requests": len(requests),
            "steps": self.steps,
            "req_per_s": len(requests) / span if span > 0 else math.inf,
            "p50_ms": 1e3 * _percentile(lat, 50),
            "p99_ms": 1e3 * _percentile(lat, 99),
            "mean_ms": 1e3 * statistics.mean(lat) if lat else 0.0,
            "span_s": span,
            "peak_working_set_bytes": self.peak_used,
            "hbm_budget_bytes": self.budget,
        }


def _percentile(sorted_vals, q) -> float:
    if not sorted_vals:
        return 0.0
    pos = (len(sorted_vals) - 1) * q / 100.0
    lo, hi = math.floor(pos), math.ceil(pos)
    return sorted_vals[lo] + (sorted_vals[hi] - sorted_vals[lo]) * (pos - lo)


def reset(requests: list[Request]) -> None:
    """Clears the timing and outputs of a trace so it can be served again."""
    for r in requests:
        r.admitted = r.launched = r.done = math.nan
        r.outputs = r.block = r.block_range = None
