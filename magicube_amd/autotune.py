"""Autotuner for the SpMM tile configuration and the fused-attention grid.

The SpMM kernel family has a fixed table of tile configurations (lanes'
gather width, column tiles per workgroup, waves per row, gathers in flight;
``spmm.tile_configs()``).  :func:`tune_spmm` times every configuration
the shape supports on the caller's own resident operands and keeps the
fastest:

* each candidate runs ``launches`` back-to-back launches captured in a HIP
  graph on one stream, replayed ``reps`` times between HIP events; the
  median time per launch is compared;
* each candidate's output is compared bitwise with the default
  configuration's, so a tuned choice never changes results;
* the result is cached on disk (JSON, ``$MC_TUNE_CACHE`` or
  ``~/.cache/magicube_amd/tune.json``) under a key of device architecture,
  pair, signedness, V, N and log2 buckets of m_vec and of the mean nonzero
  vectors per row;
* it is installed on the tuned operand (``DeviceSpmm.configs[N]``), whose
  products with that N then name it in every call; the C entry points keep
  no state, so other callers in the process are unaffected.

:func:`tune_attention` picks the persistent grid size of the fused
attention launch the same way.  The reference has no counterpart: it fixes
one tile per template (SpMM/SpMM/src/wmma_spmm.cu:2137-2153).
"""
from __future__ import annotations

import json
import math
import os
import statistics
from dataclasses import dataclass, field

import torch

from . import _lib, spmm

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "magicube_amd", "tune.json")


def cache_path() -> str:
    return os.environ.get("MC_TUNE_CACHE", DEFAULT_CACHE)


def load_cache(path: str | None = None) -> dict:
    try:
        with open(path or cache_path()) as f:
            data = json.load(f)
        return data if isinstance(data, dict) else {}
    except (OSError, ValueError):
        return {}


def save_cache(entries: dict, path: str | None = None) -> None:
    path = path or cache_path()
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(entries, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def _bucket(x: float) -> int:
    return int(math.floor(math.log2(x))) if x >= 1 else 0


def _arch(device) -> str:
    try:
        return torch.cuda.get_device_properties(device).gcnArchName.split(":")[0]
    except (AssertionError, RuntimeError, AttributeError):
        return "unknown"


def spmm_key(dev: spmm.DeviceSpmm, n: int, sign: int, arch: str) -> str:
    p = dev.prep
    mean_nnz = p.nnz_vec / p.m_vec if p.m_vec else 0.0
    return (f"spmm|{arch}|L{p.lhs_bits}R{p.rhs_bits}|{'u' if sign == _lib.MC_UNSIGNED else 's'}|V{p.V}|N{n}"
            f"|m{_bucket(p.m_vec)}|nnz{_bucket(mean_nnz)}")


@dataclass
class TuneResult:
    key: str
    config: int
    us: float
    timings: dict = field(default_factory=dict)     # config -> us per launch
    cached: bool = False


def _time(run, stream, launches: int, reps: int) -> float:
    """Median microseconds per launch over graph replays of `launches` calls."""
    with torch.cuda.stream(stream):
        run()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=stream):
        for _ in range(launches):
            run()
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    times = []
    with torch.cuda.stream(stream):
        g.replay()
        for _ in range(reps):
            t0.record(stream)
            g.replay()
            t1.record(stream)
            t1.synchronize()
            times.append(t0.elapsed_time(t1) * 1e3 / launches)
    return statistics.median(times)


def tune_spmm(dev: spmm.DeviceSpmm, rhs: torch.Tensor, n: int, sign: int = spmm.SIGNED, launches: int = 20,
              reps: int = 5, use_cache: bool = True, install: bool = True, cache_file: str | None = None) -> TuneResult:
    p = dev.prep
    key = spmm_key(dev, n, sign, _arch(rhs.device))
    entries = load_cache(cache_file) if use_cache else {}
    hit = entries.get(key)
    if hit is not None and isinstance(hit.get("config"), int):
        res = TuneResult(key, hit["config"], float(hit.get("us", 0.0)), hit.get("timings", {}), cached=True)
    else:
        stream = torch.cuda.Stream(device=rhs.device)
        m = p.m_vec * p.V
        ref = torch.empty((m, n), dtype=torch.int32, device=rhs.device)
        out = torch.empty_like(ref)
        with torch.cuda.stream(stream):
            dev.matmul(rhs, n, out=ref, sign=sign, stream=stream)          # the table default
        timings = {}
        for cfg in range(len(spmm.tile_configs())):
            try:
                with torch.cuda.stream(stream):
                    dev.matmul(rhs, n, out=out, sign=sign, config=cfg, stream=stream)
            except _lib.MagicubeError as e:
                if e.status in (_lib.MC_EUNSUPPORTED, _lib.MC_EINVAL_SHAPE):
                    continue                                           # not built / shape not divisible
                raise
            torch.cuda.synchronize()
            if not torch.equal(out, ref):
                raise RuntimeError(f"tile config {cfg} disagrees with the default configuration")
            timings[cfg] = _time(lambda c=cfg: dev.matmul(rhs, n, out=out, sign=sign, config=c, stream=stream),
                                 stream, launches, reps)
        best = min(timings, key=timings.get)
        res = TuneResult(key, best, timings[best], timings)
        if use_cache:
            entries[key] = {"config": best, "us": timings[best], "timings": {str(k): v for k, v in timings.items()}}
            save_cache(entries, cache_file)
    if install:
        dev.configs[n] = res.config
    return res


def tune_attention(batch, candidates=(0, 2048, 4096, 8192, 16384), launches: int = 3, reps: int = 3,
                   stream: torch.cuda.Stream | None = None) -> tuple[int, dict]:
    """Fastest grid size (waves; 0 = one wave per item) of the fused
    attention launch for ``batch`` (an AttentionBatch or BlockBatch)."""
    stream = stream or torch.cuda.Stream()
    timings = {gw: _time(lambda g=gw: batch.forward(stream, g), stream, launches, reps) for gw in candidates}
    return min(timings, key=timings.get), timings
