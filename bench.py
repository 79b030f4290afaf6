#!/usr/bin/env python3
"""Benchmark driver: quantized vector-sparse SpMM on one MI355X.

Workload (one "step" = one SpMM launch over the full matrix, operands
already resident in HBM): L8-R8 SpMM, A 4096 x 4096 vector-sparse with V=8,
90% sparsity (Poisson nonzero count per 8-row block, seeded), signed
full-range int8 values, dense B 4096 x 512 int8, int32 output C = A B.

The K timed steps are captured once into a HIP graph on the launch stream.
Untimed replays of it run first: at least W steps and at least
``--warmup-s`` seconds (default 0.3 s) back to back, as the GPU reaches its
working clocks only under sustained load.  Five timed replays of exactly K
steps follow, each between two torch.cuda.synchronize() calls and bracketed
by HIP events on the launch stream; ``ms_per_step`` is the median replay
divided by K.  TOPS follow the reference's formula 2 * nnz_vec * V * N / t on
unpadded nonzeros (SpMM/SpMM/spmm_benchmark.cpp:112,637).  The roofline entry
uses the compulsory bytes nnz_vec*(V + 4) + 4*(m_vec+1) + K*N + 4*M*N
against MI355X's 8 TB/s HBM3E.  The CPU baseline times the restated host
loop (compute_ref_integers) on the same arrays, single-threaded and with
OpenMP, and on the reference's 512 x 512 CPU case, on this machine's host
cores.

Prints one JSON line (under 2 KB: the headline, its roofline, the CPU
baseline and a few figures of the other sections); every section's full
results go to the ``--sections-out`` file.  ``--gpus N`` with N > 1 exits:
the benchmark measures one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0
# untimed launches before each later section's timing (the GPU is warm by then)
SECTION_WARMUP = 10
MFMA_INT8_PEAK_TOPS = 5000.0
REFERENCE_A100_TOPS = 16.67   # plot/examples/spmm_pres.txt:531-532 (DLMC 256x2304 rows, V=8, s=0.9, N=512)

# The reference's precision sweep on A100 at its published shape (DLMC rn50
# random_pruning 0.9 bottleneck_2_block_group3_5_1: 256 vector rows, K = 2304,
# N = 512, unsigned lanes): (lhs, rhs, V) -> (TOPS, plot/examples/spmm_pres.txt line).
# L12-R8 has no published figure.
PUBLISHED_SPMM_PRES = {
    (16, 16, 2): (2.26475, 110), (16, 16, 4): (4.46368, 121), (16, 16, 8): (10.2298, 132),
    (16, 8, 2): (3.64078, 310), (16, 8, 4): (7.03666, 321), (16, 8, 8): (15.7426, 332),
    (8, 8, 2): (4.34606, 510), (8, 8, 4): (8.38464, 521), (8, 8, 8): (16.6716, 532),
    (16, 4, 2): (4.29697, 710), (16, 4, 4): (8.23018, 721), (16, 4, 8): (17.4599, 732),
    (12, 4, 2): (4.25792, 910), (12, 4, 4): (8.47903, 921), (12, 4, 8): (20.1014, 932),
    (8, 4, 2): (4.63334, 1110), (8, 4, 4): (8.65352, 1121), (8, 4, 8): (20.4712, 1132),
    (4, 4, 2): (5.95672, 1310), (4, 4, 4): (11.6777, 1321), (4, 4, 8): (21.0431, 1332),
}
# The reference's SpMM ablation on A100 (DLMC rn50 random_pruning
# bottleneck_2_block_group3_5_1: 256 vector rows, K = 2304, N = 512):
# (stage, lhs, rhs, sparsity, V) -> (TOPS, plot/examples/spmm_abl_study.txt line).
# Stages: basic kernel, + conflict-free shared memory, + prefetch, + shuffle
# (the last for 4-bit B only).
PUBLISHED_SPMM_ABLATION = {
    ("basic", 16, 8, 0.7, 2): (3.7567, 12), ("basic", 16, 8, 0.7, 8): (15.0338, 24),
    ("basic", 16, 8, 0.9, 2): (2.74522, 36), ("basic", 16, 8, 0.9, 8): (11.5881, 48),
    ("basic", 8, 8, 0.7, 2): (3.85332, 62), ("basic", 8, 8, 0.7, 8): (15.3045, 74),
    ("basic", 8, 8, 0.9, 2): (2.97658, 86), ("basic", 8, 8, 0.9, 8): (11.7619, 98),
    ("basic", 8, 4, 0.7, 2): (4.55222, 112), ("basic", 8, 4, 0.7, 8): (18.5377, 123),
    ("basic", 8, 4, 0.9, 2): (3.17841, 134), ("basic", 8, 4, 0.9, 8): (13.1074, 145),
    ("basic", 4, 4, 0.7, 2): (4.93733, 159), ("basic", 4, 4, 0.7, 8): (19.0693, 171),
    ("basic", 4, 4, 0.9, 2): (3.62971, 183), ("basic", 4, 4, 0.9, 8): (13.4169, 195),
    ("conflict_free", 16, 8, 0.7, 2): (4.37232, 212), ("conflict_free", 16, 8, 0.7, 8): (17.4843, 224),
    ("conflict_free", 16, 8, 0.9, 2): (3.03427, 236), ("conflict_free", 16, 8, 0.9, 8): (13.0248, 248),
    ("conflict_free", 8, 8, 0.7, 2): (4.47763, 265), ("conflict_free", 8, 8, 0.7, 8): (17.8471, 277),
    ("conflict_free", 8, 8, 0.9, 2): (3.41112, 289), ("conflict_free", 8, 8, 0.9, 8): (13.2815, 301),
    ("conflict_free", 8, 4, 0.7, 2): (4.7151, 318), ("conflict_free", 8, 4, 0.7, 8): (18.9403, 329),
    ("conflict_free", 8, 4, 0.9, 2): (3.25536, 340), ("conflict_free", 8, 4, 0.9, 8): (13.3273, 351),
    ("conflict_free", 4, 4, 0.7, 2): (4.95606, 368), ("conflict_free", 4, 4, 0.7, 8): (19.4581, 380),
    ("conflict_free", 4, 4, 0.9, 2): (3.65532, 392), ("conflict_free", 4, 4, 0.9, 8): (13.6976, 404),
    ("prefetch", 16, 8, 0.7, 2): (5.53128, 421), ("prefetch", 16, 8, 0.7, 8): (21.4967, 433),
    ("prefetch", 16, 8, 0.9, 2): (3.6021, 445), ("prefetch", 16, 8, 0.9, 8): (15.7921, 457),
    ("prefetch", 8, 8, 0.7, 2): (5.86861, 474), ("prefetch", 8, 8, 0.7, 8): (22.9959, 486),
    ("prefetch", 8, 8, 0.9, 2): (4.32687, 498), ("prefetch", 8, 8, 0.9, 8): (16.5091, 510),
    ("prefetch", 8, 4, 0.7, 2): (6.23654, 527), ("prefetch", 8, 4, 0.7, 8): (24.4017, 538),
    ("prefetch", 8, 4, 0.9, 2): (4.02061, 549), ("prefetch", 8, 4, 0.9, 8): (16.6622, 560),
    ("prefetch", 4, 4, 0.7, 2): (6.65973, 577), ("prefetch", 4, 4, 0.7, 8): (25.4291, 589),
    ("prefetch", 4, 4, 0.9, 2): (4.6343, 601), ("prefetch", 4, 4, 0.9, 8): (17.2339, 613),
    ("shuffle", 8, 4, 0.7, 2): (8.49537, 630), ("shuffle", 8, 4, 0.7, 8): (34.7546, 642),
    ("shuffle", 8, 4, 0.9, 2): (4.64603, 654), ("shuffle", 8, 4, 0.9, 8): (20.5391, 666),
    ("shuffle", 4, 4, 0.7, 2): (9.52411, 684), ("shuffle", 4, 4, 0.7, 8): (36.9262, 697),
    ("shuffle", 4, 4, 0.9, 2): (5.9758, 710), ("shuffle", 4, 4, 0.9, 8): (21.0911, 723),
}
# The reference's A100 runs on DLMC rn50 random_pruning bottleneck_2_block_group3_5_1
# (256 vector rows, K = 2304), L8-R8 Magicube vs the int8 baselines, per-launch event timing:
# (sparsity, V, N) -> (Magicube ms, cuBLAS int8 dense ms, cuSPARSE int8 Blocked-ELL ms,
#                      and their lines in plot/examples/spmm_magicube_8b8b.txt,
#                      spmm_cublas_int8.txt, spmm_cusparse_int8.txt).
PUBLISHED_DLMC_A100 = {
    (0.5, 2, 128): (0.0587884, 0.0362348, 0.132606, 539, 449, 539),
    (0.7, 2, 128): (0.0394848, 0.035898, 0.0844892, 2075, 1729, 2075),
    (0.8, 2, 128): (0.0322237, 0.0369709, 0.0625794, 3611, 3009, 3611),
    (0.9, 2, 128): (0.0216988, 0.0356188, 0.0411213, 5147, 4289, 5147),
    (0.95, 2, 128): (0.0158976, 0.0360255, 0.0285692, 6683, 5569, 6683),
    (0.98, 2, 128): (0.0125451, 0.0380169, 0.0241786, 8219, 6849, 8219),
    (0.5, 4, 128): (0.0619648, 0.111239, 0.132816, 9755, 8129, 9755),
    (0.7, 4, 128): (0.0414849, 0.110935, 0.088539, 11291, 9409, 11291),
    (0.8, 4, 128): (0.0311731, 0.111121, 0.0647134, 12827, 10689, 12827),
    (0.9, 4, 128): (0.0208251, 0.111213, 0.0407606, 14363, 11969, 14363),
    (0.95, 4, 128): (0.0158529, 0.111089, 0.0278881, 15899, 13249, 15899),
    (0.98, 4, 128): (0.0132805, 0.112829, 0.024652, 17435, 14529, 17435),
    (0.5, 8, 128): (0.0581013, 0.177379, 0.0536916, 18971, 15809, 18971),
    (0.7, 8, 128): (0.0399071, 0.177774, 0.0363102, 20507, 17089, 20507),
    (0.8, 8, 128): (0.0301272, 0.178686, 0.0267632, 22043, 18369, 22043),
    (0.9, 8, 128): (0.0204869, 0.177367, 0.0205262, 23579, 19649, 23579),
    (0.95, 8, 128): (0.0155595, 0.180883, 0.0165044, 25115, 20929, 25115),
    (0.98, 8, 128): (0.0133504, 0.179176, 0.0150234, 26651, 22209, 26651),
    (0.5, 2, 256): (0.0631769, 0.113565, 0.196471, 28187, 23489, 28187),
    (0.7, 2, 256): (0.0421415, 0.113455, 0.16037, 29723, 24769, 29723),
    (0.8, 2, 256): (0.0334617, 0.1136, 0.108045, 31259, 26049, 31259),
    (0.9, 2, 256): (0.0241716, 0.1161, 0.0721037, 32795, 27329, 32795),
    (0.95, 2, 256): (0.0165831, 0.113693, 0.0469088, 34331, 28609, 34331),
    (0.98, 2, 256): (0.0137907, 0.115026, 0.0397939, 35867, 29889, 35867),
    (0.5, 4, 256): (0.0662818, 0.178334, 0.196769, 37403, 31169, 37403),
    (0.7, 4, 256): (0.0461313, 0.178086, 0.132066, 38939, 32449, 38939),
    (0.8, 4, 256): (0.0353168, 0.179439, 0.116964, 40475, 33729, 40475),
    (0.9, 4, 256): (0.0227795, 0.178115, 0.0755144, 42011, 35009, 42011),
    (0.95, 4, 256): (0.019352, 0.181551, 0.0472593, 43547, 36289, 43547),
    (0.98, 4, 256): (0.0141384, 0.178807, 0.0401977, 45083, 37569, 45083),
    (0.5, 8, 256): (0.0631796, 0.116912, 0.0885132, 46619, 38849, 46619),
    (0.7, 8, 256): (0.0430592, 0.11651, 0.0569483, 48155, 40129, 48155),
    (0.8, 8, 256): (0.0357006, 0.115956, 0.0432874, 49691, 41409, 49691),
    (0.9, 8, 256): (0.0226486, 0.117112, 0.0309084, 51227, 42689, 51227),
    (0.95, 8, 256): (0.0173295, 0.116243, 0.0227704, 52763, 43969, 52763),
    (0.98, 8, 256): (0.0147643, 0.1168, 0.0216203, 54299, 45249, 54299),
}
# SDDMM on A100 at the same matrix, N = 2304, K = 512 (basic kernel):
# (bits, V) -> (TOPS, plot/examples/sddmm_abl_study.txt line)
PUBLISHED_SDDMM_K512 = {
    (16, 2): (2.23103, 58), (16, 4): (4.58991, 64), (16, 8): (9.71052, 70),
    (8, 2): (3.32038, 166), (8, 4): (6.5755, 172), (8, 8): (12.6419, 178),
    (4, 2): (4.44799, 284), (4, 4): (8.83131, 291), (4, 8): (17.2533, 298),
}


def make_workload(M, K, N, V, sparsity, seed, lhs_bits=8, rhs_bits=8, signed=True):
    from magicube_amd import formats, synthetic
    csr = synthetic.poisson_block_mask(M, K, V, sparsity, seed)
    a = synthetic.random_lanes(csr.nnz_vec * V, lhs_bits, signed, seed + 1)
    b = synthetic.random_lanes(K * N, rhs_bits, signed, seed + 2)
    a_packed = formats.pack_lanes(a, lhs_bits)
    prep = formats.prepare_spmm(csr, a_packed, lhs_bits, rhs_bits)
    rhs = formats.pack_lanes(b, rhs_bits).view(np.int32)
    return csr, a, b, a_packed, prep, rhs


def algorithmic_bytes(nnz_vec, m_vec, M, K, N, V, lhs_bits=8, rhs_bits=8):
    cb = 16 if lhs_bits == 12 else lhs_bits       # 12-bit lanes are stored in 16-bit containers
    return nnz_vec * (V * cb // 8 + 4) + 4 * (m_vec + 1) + K * N * rhs_bits // 8 + 4 * M * N


def checksum_ok(C, csr, a, b, K, N, V, seed=0):
    rng = np.random.default_rng(seed)
    w = rng.integers(0, 1 << 10, size=N)
    Bw = b.astype(np.int64).reshape(K, N) @ w
    av = a.astype(np.int64).reshape(-1, V)
    rows = np.repeat(np.arange(csr.m_vec), np.diff(csr.row_offsets))
    exp = np.zeros((csr.m_vec, V), np.int64)
    np.add.at(exp, rows, av * Bw[csr.col_indices][:, None])
    got = C.astype(np.int64) @ w
    return bool(np.array_equal(got & 0xFFFFFFFF, exp.reshape(-1) & 0xFFFFFFFF))


def time_gpu(run_once, steps, warmup, stream, warm_s=0.02, detail=False):
    """Per-step time (ms) of ``steps`` graph-replayed launches.

    A few eager launches initialize the work; the ``steps`` launches are then
    captured once into a HIP graph, the graph is replayed untimed until
    ``warmup`` steps and at least ``warm_s`` seconds of back-to-back work
    have run (the GPU reaches its working clocks only under sustained load,
    not between eager launches or host-side setup), and five timed replays
    follow, each bracketed by torch.cuda.synchronize() and by HIP events on
    the capture stream.  Returns the median over the five and the graph; with
    ``detail`` a dict instead: that median as ``ms`` (every timed replay is
    exactly ``steps`` launches), the first replay alone (which runs right
    after the host has waited on the GPU and is 2-3 % slower), the host wall
    time of the median replay and the warm-up actually run."""
    import torch
    eager = min(warmup, 3)
    for _ in range(eager):
        run_once()
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=stream):
        for _ in range(steps):
            run_once()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    reps = 5
    times, walls = [], []
    # replay() launches on the current stream: keep it the capture stream so
    # the events bracket exactly the replayed launches
    with torch.cuda.stream(stream):
        t0 = time.perf_counter()
        done = 0
        while done < max(1, -(-(warmup - eager) // steps)) or time.perf_counter() - t0 < warm_s:
            graph.replay()      # untimed
            done += 1
            if done % 8 == 0:
                torch.cuda.synchronize()
        torch.cuda.synchronize()
        warm_wall = time.perf_counter() - t0
        for _ in range(reps):
            torch.cuda.synchronize()
            h0 = time.perf_counter()
            start.record(stream)
            graph.replay()
            end.record(stream)
            torch.cuda.synchronize()
            walls.append(time.perf_counter() - h0)
            times.append(start.elapsed_time(end) / steps)
    if not detail:
        return statistics.median(times), graph
    order = sorted(range(reps), key=lambda i: times[i])
    mid = order[reps // 2]
    return {"ms": times[mid], "first_ms": times[0], "min_ms": min(times), "max_ms": max(times),
            "host_wall_ms": walls[mid] * 1e3, "warmup_steps": eager + done * steps,
            "warmup_s": warm_wall}, graph


def host_cpu():
    """(model, threads this process may use) of the host.  The model is the
    first processor's model name, its family/model/stepping numbers, the
    logical CPUs the host shows and AVX-512 when present (a virtual machine
    may report only a generic name)."""
    fields = {}
    n_cpu = 0
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if ":" not in line:
                    continue
                k, v = (x.strip() for x in line.split(":", 1))
                if k == "processor":
                    n_cpu += 1
                if n_cpu == 1 and k in ("model name", "cpu family", "model", "stepping", "flags"):
                    fields[k] = v
    except OSError:
        pass
    model = fields.get("model name", "")
    if model:
        ids = "/".join(fields.get(k, "?") for k in ("cpu family", "model", "stepping"))
        avx = ", AVX-512" if "avx512f" in fields.get("flags", "").split() else ""
        model = f"{model}, {ids}, {n_cpu} CPUs{avx}"
    try:
        threads = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        threads = os.cpu_count() or 1
    # the process's share of a shared host, if the environment states it
    cap = os.environ.get("OMP_NUM_THREADS", "")
    if cap.isdigit() and int(cap) > 0:
        threads = min(threads, int(cap))
    return model, threads


def config1_workload(seed=5):
    """The reference's CPU-runnable case: A 512 x 512 vector-sparse, V = 8,
    round(0.1 * 512) = 51 column vectors per 8-row block drawn uniformly
    without replacement, int8 values uniform in [-128, 127]; B 512 x 128 int8
    uniform; C = A B in int32."""
    from magicube_amd import formats, synthetic
    M = K = 512
    N, V = 128, 8
    csr = synthetic.uniform_block_mask(M, K, V, 0.9, seed)
    a = synthetic.random_lanes(csr.nnz_vec * V, 8, True, seed + 1)
    b = synthetic.random_lanes(K * N, 8, True, seed + 2)
    a_packed = formats.pack_lanes(a, 8)
    return csr, a, b, a_packed, formats.pack_lanes(b, 8).view(np.int32), (M, K, N, V)


def _time_host(fn, min_runs, budget_s, max_runs=30):
    fn()                                            # warm-up
    ts = []
    t_end = time.time() + budget_s
    while len(ts) < min_runs or (time.time() < t_end and len(ts) < max_runs):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return statistics.median(ts), len(ts)


def cpu_baseline(csr, a_packed, rhs, M, K, N, V, budget_s=10.0):
    """The reference's host loop compute_ref_integers
    (SpMM/SpMM/spmm_benchmark.cpp:69-120), restated in oracle/ (checker
    only), timed on this host: single-threaded (the reference's own CPU
    path) on the headline arrays, then row-parallel with OpenMP on every
    thread this process may use, whose C must equal the single-thread C, and
    single-threaded on the reference's 512 x 512, N = 128 case, whose C must
    equal the GPU's.  The library built for x86-64-v4 is used where the CPU
    has AVX-512 (the reference builds -march=native)."""
    import torch
    import oracle
    from magicube_amd import formats, spmm
    handle, march = oracle.host_lib()
    model, threads = host_cpu()
    A = np.ascontiguousarray(a_packed).view(np.int64)
    flops = 2.0 * csr.nnz_vec * V * N
    ref = {}

    def single():
        ref["C"], _ = oracle.spmm_ref_check(A, 8, rhs, M, K, N, 8, 8, 8, V, csr.row_offsets, csr.col_indices,
                                            csr.m_vec, 1, handle=handle)

    st_s, st_n = _time_host(single, 5, budget_s)

    par = {}

    def parallel():
        par["C"], _ = oracle.spmm_ref_check_omp(A, 8, rhs, M, K, N, 8, 8, V, csr.row_offsets, csr.col_indices,
                                                csr.m_vec, threads, handle=handle)

    mt_s, mt_n = _time_host(parallel, 5, 3.0)
    par_ok = bool(np.array_equal(par["C"], ref["C"]))

    # the reference's CPU-runnable case (512 x 512, V = 8, N = 128), on the GPU too
    c1, a1, b1, ap1, rhs1, (M1, K1, N1, V1) = config1_workload()
    A1 = np.ascontiguousarray(ap1).view(np.int64)
    c1_ref = {}

    def single1():
        c1_ref["C"], _ = oracle.spmm_ref_check(A1, 8, rhs1, M1, K1, N1, 8, 8, 8, V1, c1.row_offsets,
                                               c1.col_indices, c1.m_vec, 1, handle=handle)

    s1_s, s1_n = _time_host(single1, 20, 2.0, max_runs=2000)
    prep1 = formats.prepare_spmm(c1, ap1, 8, 8)
    dev1 = spmm.DeviceSpmm.upload(prep1)
    out1 = dev1.matmul(torch.from_numpy(rhs1).cuda(), N1, sign=spmm.SIGNED)
    torch.cuda.synchronize()
    # the literal check reads A lanes unsigned and sign-extends the top B
    # lane (spmm_benchmark.cpp:101,106), so on signed full-range data the GPU
    # is compared with the plain signed sum
    plain1 = oracle.spmm_plain(c1.m_vec, V1, N1, c1.row_offsets, c1.col_indices, a1, b1)
    c1_gpu_exact = bool(np.array_equal(out1.cpu().numpy(), plain1))
    flops1 = 2.0 * c1.nnz_vec * V1 * N1
    return {
        "value": flops / st_s / 1e12,
        "unit": "TOPS",
        "cores": 1,
        "kind": "port",
        "sample": f"headline arrays, compute_ref_integers 1 thread, median of {st_n} x {st_s * 1e3:.0f} ms, "
                  f"-O3 -march={march}",
        "verified": par_ok,
        "cpu_model": model,
        "parallel": {"value": flops / mt_s / 1e12, "cores": threads, "ms": mt_s * 1e3, "runs": mt_n,
                     "kind": "port-openmp", "equals_single_thread": par_ok},
        "config1": {"workload": "512x512 V=8 N=128 int8", "nnz_vec": c1.nnz_vec,
                    "value": flops1 / s1_s / 1e12, "ms": s1_s * 1e3, "runs": s1_n, "cores": 1,
                    "gpu_equals_plain_sum": c1_gpu_exact},
    }


def sddmm_bench(steps, warmup, seed=31):
    """SDDMM at M = N = 4096, V = 8, K in {64, 128}, sparsity in {0.9, 0.95},
    signed int8 and int4, reference output layout; exact check on sampled
    vector rows; TOPS on unpadded nonzeros (2 * nnz_vec * V * K) and the
    HBM roofline on compulsory bytes M*K*b/8 + N*K*b/8 + 4*nnz_vec +
    4*(m_vec + 1) + 4*V*nnz_vec."""
    import torch
    from magicube_amd import formats, sddmm, synthetic
    stream = torch.cuda.Stream()
    M = N = 4096
    V = 8
    out = []
    for bits in (8, 4):
        for K in (64, 128):
            for s in (0.9, 0.95):
                csr = synthetic.uniform_block_mask(M, N, V, s, seed)
                mask = sddmm.DeviceMask.upload(csr)
                a = synthetic.random_lanes(M * K, bits, True, seed + 1)
                b = synthetic.random_lanes(N * K, bits, True, seed + 2)
                lhs = torch.from_numpy(formats.pack_lanes(a, bits).view(np.int32)).cuda()
                rhs = torch.from_numpy(formats.pack_lanes(b, bits).view(np.int32)).cuda()
                res = torch.zeros(mask.output_size(), dtype=torch.int32, device="cuda")

                def run():
                    mask.sddmm(bits, lhs, rhs, K, out=res, stream=stream)

                with torch.cuda.stream(stream):
                    run()
                torch.cuda.synchronize()
                got = res.cpu().numpy()
                # exact check of 16 sampled vector rows
                A = a.astype(np.int64).reshape(M, K)
                B = b.astype(np.int64).reshape(N, K)
                ok = True
                for i in np.random.default_rng(seed).choice(csr.m_vec, 16, replace=False):
                    s0, e0 = mask.pairs[2 * i], mask.pairs[2 * i + 1]
                    cols = csr.col_indices[csr.row_offsets[i]:csr.row_offsets[i + 1]]
                    exp = (A[i * V:(i + 1) * V] @ B[cols].T) & 0xFFFFFFFF
                    j = np.arange(s0, e0)
                    idx = (j[None, :] // 8) * 8 * V + 8 * np.arange(V)[:, None] + (j[None, :] % 8)
                    ok &= bool(np.array_equal(got[idx].astype(np.int64) & 0xFFFFFFFF, exp))
                ms, _ = time_gpu(run, steps, warmup, stream)
                nnz = csr.nnz_vec
                alg = M * K * bits // 8 + N * K * bits // 8 + 4 * nnz + 4 * (csr.m_vec + 1) + 4 * V * nnz
                out.append({"bits": bits, "K": K, "sparsity": s, "nnz_vec": nnz, "us": ms * 1e3,
                            "tops": 2.0 * nnz * V * K / (ms * 1e-3) / 1e12,
                            "hbm_frac": alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "verified": ok})
    return {"workload": "SDDMM M=N=4096, V=8, uniform mask, signed int8/int4, reference output layout",
            "cases": out}


def precision_bench(steps, warmup, seed=17):
    """Every precision pair and V at the reference's published sweep shape
    (256 vector rows, K = 2304, N = 512, sparsity 0.9, unsigned lanes as the
    reference's MMAs), beside the A100 figure; checksum-verified; default
    tile configuration."""
    import torch
    from magicube_amd import spmm
    stream = torch.cuda.Stream()
    K, N, s = 2304, 512, 0.9
    cases = []
    for (la, lb) in spmm.PAIRS:
        for V in (2, 4, 8):
            M = 256 * V
            csr, a, b, _, prep, rhs_np = make_workload(M, K, N, V, s, seed, la, lb, signed=False)
            dev = spmm.DeviceSpmm.upload(prep)
            rhs = torch.from_numpy(rhs_np).cuda()
            out = torch.empty((M, N), dtype=torch.int32, device="cuda")

            def run():
                dev.matmul(rhs, N, out=out, sign=spmm.UNSIGNED, stream=stream)

            with torch.cuda.stream(stream):
                run()
            torch.cuda.synchronize()
            ok = checksum_ok(out.cpu().numpy(), csr, a, b, K, N, V)
            ms, _ = time_gpu(run, steps, warmup, stream)
            nnz = csr.nnz_vec
            tops = 2.0 * nnz * V * N / (ms * 1e-3) / 1e12
            alg = algorithmic_bytes(nnz, csr.m_vec, M, K, N, V, la, lb)
            pub = PUBLISHED_SPMM_PRES.get((la, lb, V))
            cases.append({"pair": f"L{la}-R{lb}", "V": V, "nnz_vec": nnz, "us": ms * 1e3, "tops": tops,
                          "hbm_frac": alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                          "a100_tops": pub[0] if pub else None,
                          "ratio_vs_a100": tops / pub[0] if pub else None, "verified": ok})
    return {"workload": "SpMM m_vec=256, K=2304, N=512, s=0.9 (Poisson vector-sparse), unsigned, every pair x V",
            "a100_source": "plot/examples/spmm_pres.txt", "cases": cases}


def ablation_bench(steps, warmup, seed=23):
    """SpMM ablation at the reference's ablation shape (256 vector rows,
    K = 2304, N = 512, s in {0.7, 0.9}, V in {2, 8}, unsigned; pairs L16-R8,
    L8-R8, L8-R4, L4-R4), beside its A100 stages.  The stages here are this
    kernel's own design steps: gathers of one, two or three 64-slot k-blocks in
    flight per wave (tile configs 13, 14, 7), the per-pair default tile, and
    the default with vector rows in their stored order instead of sorted by
    length."""
    import torch
    from magicube_amd import spmm
    stream = torch.cuda.Stream()
    K, N = 2304, 512
    stages = (("one_kblock_in_flight", 13, "sorted"), ("two_kblocks_in_flight", 14, "sorted"),
              ("three_kblocks_in_flight", 7, "sorted"), ("default", None, "sorted"),
              ("default_rows_unsorted", None, "identity"))
    cases = []
    for (la, lb) in ((16, 8), (8, 8), (8, 4), (4, 4)):
        for s in (0.7, 0.9):
            for V in (2, 8):
                M = 256 * V
                csr, a, b, _, prep, rhs_np = make_workload(M, K, N, V, s, seed, la, lb, signed=False)
                dev = spmm.DeviceSpmm.upload(prep)
                rhs = torch.from_numpy(rhs_np).cuda()
                out = torch.empty((M, N), dtype=torch.int32, device="cuda")
                sorted_rows = dev.row_indices
                identity = torch.arange(csr.m_vec, dtype=torch.int32, device="cuda")
                row = {"pair": f"L{la}-R{lb}", "sparsity": s, "V": V, "nnz_vec": csr.nnz_vec, "tops": {},
                       "verified": True}
                for name, cfg, order in stages:
                    dev.row_indices = sorted_rows if order == "sorted" else identity

                    def run():
                        dev.matmul(rhs, N, out=out, sign=spmm.UNSIGNED, config=cfg, stream=stream)

                    with torch.cuda.stream(stream):
                        run()
                    torch.cuda.synchronize()
                    row["verified"] &= checksum_ok(out.cpu().numpy(), csr, a, b, K, N, V)
                    ms, _ = time_gpu(run, steps, warmup, stream)
                    row["tops"][name] = 2.0 * csr.nnz_vec * V * N / (ms * 1e-3) / 1e12
                row["a100_tops"] = {st: PUBLISHED_SPMM_ABLATION[(st, la, lb, s, V)][0]
                                    for st in ("basic", "conflict_free", "prefetch", "shuffle")
                                    if (st, la, lb, s, V) in PUBLISHED_SPMM_ABLATION}
                cases.append(row)
    return {"workload": "SpMM m_vec=256, K=2304, N=512, s in {0.7, 0.9}, V in {2, 8}, unsigned (Poisson mask)",
            "a100_source": "plot/examples/spmm_abl_study.txt", "cases": cases}


def sddmm_published_bench(steps, warmup, seed=43):
    """SDDMM at the reference's published configuration (256 vector rows,
    N = 2304, K = 512, sparsity 0.9, unsigned lanes), every width and V,
    beside the A100 figure; exact check on sampled vector rows."""
    import torch
    from magicube_amd import formats, sddmm, synthetic
    stream = torch.cuda.Stream()
    N, K, s = 2304, 512, 0.9
    cases = []
    for bits in (16, 8, 4):
        for V in (2, 4, 8):
            M = 256 * V
            csr = synthetic.poisson_block_mask(M, N, V, s, seed)
            mask = sddmm.DeviceMask.upload(csr)
            a = synthetic.random_lanes(M * K, bits, False, seed + 1)
            b = synthetic.random_lanes(N * K, bits, False, seed + 2)
            lhs = torch.from_numpy(formats.pack_lanes(a, bits).view(np.int32)).cuda()
            rhs = torch.from_numpy(formats.pack_lanes(b, bits).view(np.int32)).cuda()
            res = torch.zeros(mask.output_size(), dtype=torch.int32, device="cuda")

            def run():
                mask.sddmm(bits, lhs, rhs, K, out=res, sign=sddmm.UNSIGNED, stream=stream)

            with torch.cuda.stream(stream):
                run()
            torch.cuda.synchronize()
            got = res.cpu().numpy()
            A = a.astype(np.int64).reshape(M, K)
            B = b.astype(np.int64).reshape(N, K)
            ok = True
            for i in np.random.default_rng(seed).choice(csr.m_vec, 8, replace=False):
                s0, e0 = mask.pairs[2 * i], mask.pairs[2 * i + 1]
                cols = csr.col_indices[csr.row_offsets[i]:csr.row_offsets[i + 1]]
                exp = (A[i * V:(i + 1) * V] @ B[cols].T) & 0xFFFFFFFF
                j = np.arange(s0, e0)
                idx = (j[None, :] // 8) * 8 * V + 8 * np.arange(V)[:, None] + (j[None, :] % 8)
                ok &= bool(np.array_equal(got[idx].astype(np.int64) & 0xFFFFFFFF, exp))
            ms, _ = time_gpu(run, steps, warmup, stream)
            nnz = csr.nnz_vec
            tops = 2.0 * nnz * V * K / (ms * 1e-3) / 1e12
            alg = M * K * bits // 8 + N * K * bits // 8 + 4 * nnz + 4 * (csr.m_vec + 1) + 4 * V * nnz
            pub = PUBLISHED_SDDMM_K512[(bits, V)]
            cases.append({"bits": bits, "V": V, "nnz_vec": nnz, "us": ms * 1e3, "tops": tops,
                          "hbm_frac": alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "a100_tops": pub[0],
                          "ratio_vs_a100": tops / pub[0], "verified": ok})
    return {"workload": "SDDMM m_vec=256, N=2304, K=512, s=0.9 (Poisson vector-sparse), unsigned, 16/8/4-bit x V",
            "a100_source": "plot/examples/sddmm_abl_study.txt", "cases": cases}


# The reference's 4-layer encoder on A100 (embed 256, 4 heads, MLP 1024, seq
# 4096, sparsity 0.9, V = 8, fp16 model), mean of 8 runs:
# (batch, lhs, rhs) -> (ms, plot/examples/magicube_n2n.txt lines); dense from pytorch_n2n.txt.
PUBLISHED_ENCODER = {(2, 8, 8): (3.731, "19-26"), (2, 8, 4): (3.371, "10-17"),
                     (2, 16, 8): (3.754, "28-35"), (2, 4, 4): (3.338, "1-8"),
                     (8, 8, 8): (13.261, "55-62"), (8, 8, 4): (12.459, "46-53"),
                     (8, 16, 8): (13.443, "64-71"), (8, 4, 4): (12.048, "37-44")}
PUBLISHED_ENCODER_DENSE = {2: (5.947, "1-8"), 8: (24.541, "10-17")}


def encoder_bench(reps=10, warmup=3, seed=0):
    """The reference's end-to-end encoder benchmark (end_to_end.py): four
    sparse blocks at seq 4096, embed 256, 4 heads, MLP 1024, sparsity 0.9,
    V = 8, fp16 model, inference forward over random tokens; per batch size
    and (lhs, rhs) precision the eager forward (HIP events around ``reps``
    forwards) and the same forward replayed from a HIP graph; the dense
    encoder (torch multi-head attention) beside it."""
    import torch
    from magicube_amd import spattention
    cases = []
    stream = torch.cuda.Stream()

    def timed(fn):
        with torch.cuda.stream(stream):
            for _ in range(warmup):
                fn()
            t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            stream.synchronize()
            t0.record(stream)
            for _ in range(reps):
                fn()
            t1.record(stream)
            t1.synchronize()
        return t0.elapsed_time(t1) / reps

    torch.manual_seed(seed)
    with torch.no_grad():
        for bs in (2, 8):
            x = torch.randint(0, 257, (bs, 4096), dtype=torch.int32, device="cuda")
            for (lp, rp) in ((8, 8), (16, 8), (8, 4), (4, 4)):
                model = spattention.SparseTransformerEncoder(257, 256, 4, 0.9, 4096, 1024, 1024, vec_length=8,
                                                             lhs_pre=lp, rhs_pre=rp, seed=seed).cuda().half().eval()
                eager = timed(lambda: model(x))
                with torch.cuda.stream(stream):
                    g = torch.cuda.CUDAGraph()
                    with torch.cuda.graph(g, stream=stream):
                        out = model(x)
                graph = timed(g.replay)
                ok = bool(torch.isfinite(out.float()).all())
                pub = PUBLISHED_ENCODER[(bs, lp, rp)]
                cases.append({"batch": bs, "lhs_pre": lp, "rhs_pre": rp, "eager_ms": eager, "graph_ms": graph,
                              "a100_ms": pub[0], "speedup_vs_a100": pub[0] / eager, "finite": ok})
                del model, g, out
            dense = spattention.TransformerEncoder(257, 256, 4, 4096, 1024, 1024).cuda().half().eval()
            dms = timed(lambda: dense(x))
            dpub = PUBLISHED_ENCODER_DENSE[bs]
            cases.append({"batch": bs, "dense": True, "eager_ms": dms, "a100_ms": dpub[0],
                          "speedup_vs_a100": dpub[0] / dms})
            del dense
            torch.cuda.empty_cache()
    return {"workload": "4-layer encoder, seq 4096, embed 256, 4 heads, MLP 1024, sparsity 0.9, V=8, fp16, "
                        "random tokens and weights, inference forward",
            "a100_source": "plot/examples/magicube_n2n.txt, plot/examples/pytorch_n2n.txt",
            "note": "rhs_pre 4: int4 Q and K, V quantized to the 4-bit range as the reference does", "cases": cases}


def dense_attention_fp64_torch(q, k, v, mask, V):
    """softmax(Q K^T / 8) V in float64 on the GPU over the vector-sparse mask
    (rows of a vector row with no column give zeros): the FP64 reference the
    quantized layer's error is stated against."""
    import torch
    L = q.shape[0]
    keep = torch.zeros((L // V, L), dtype=torch.bool, device=q.device)
    rows = torch.repeat_interleave(torch.arange(L // V, device=q.device),
                                   torch.from_numpy(np.diff(mask.row_offsets).astype(np.int64)).to(q.device))
    keep[rows, torch.from_numpy(mask.col_indices.astype(np.int64)).to(q.device)] = True
    keep = keep.repeat_interleave(V, dim=0)
    s = (q.double() @ k.double().T) / 8.0
    s = s.masked_fill(~keep, float("-inf"))
    p = torch.softmax(s, dim=1).nan_to_num(0.0)
    return p @ v.double()


def attention_bench(steps, warmup, seed=101, samples=8, heads=16, V=8, sparsity=0.9):
    """Batched quantized sparse attention: samples x heads matrices, L ~ U[512, 4096]
    (multiple of 64) per sample, one mask per (sample, head); fused persistent
    launch in a HIP graph vs the one-op-per-launch pipeline."""
    import torch
    from magicube_amd import attention, synthetic
    rng = np.random.default_rng(seed)
    Ls = [int(x) * 64 for x in rng.integers(512 // 64, 4096 // 64 + 1, size=samples)]
    mats, masks = [], []
    for s_i, L in enumerate(Ls):
        for h in range(heads):
            sd = seed + 1000 * s_i + h
            q, k, v = synthetic.attention_inputs(L, sd)
            mask = synthetic.attention_mask(L, V, sparsity, sd + 7)
            mats.append(attention.AttentionMatrix(*(torch.from_numpy(x).cuda() for x in (q, k, v)), mask))
            masks.append(mask)
    batch = attention.AttentionBatch(mats, qk_bits=8, V=V, p_mode=attention.P_ROW_SCALED, p_scale=127.0)
    stream = torch.cuda.Stream()
    nnz = sum(m.nnz_vec for m in masks)
    ops = 2.0 * 2.0 * nnz * V * attention.HEAD_DIM
    alg = sum(3 * m.L * 64 * 2 + m.L * 64 * 4 + 4 * m.mask.nnz_vec + 4 * (m.mask.m_vec + 1) for m in mats)
    # the one-op-per-launch pipeline also writes and reads back the int8
    # Q/K/V, the int32 S, the int8 P and the int32 P V
    unfused = alg + sum(2 * 3 * m.L * 64 + 2 * 4 * V * m.mask.nnz_vec + 2 * V * m.mask.nnz_vec + 2 * 4 * m.L * 64
                        for m in mats)

    def fused():
        batch.forward(stream)

    def plain():
        batch.forward_plain(stream)

    with torch.cuda.stream(stream):
        plain()
    torch.cuda.synchronize()
    plain_out = [d.o.clone() for d in batch.mats]
    fused_ms, _ = time_gpu(fused, steps, warmup, stream)
    identical = all(torch.equal(a.view(torch.int32), d.o.view(torch.int32)) for a, d in zip(plain_out, batch.mats))
    plain_graph_ms, _ = time_gpu(plain, max(1, steps // 4), 2, stream)
    # plain pipeline launched eagerly, as its callers do (launch overhead included)
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    eager = []
    with torch.cuda.stream(stream):
        for _ in range(5):
            torch.cuda.synchronize()
            start.record(stream)
            plain()
            end.record(stream)
            torch.cuda.synchronize()
            eager.append(start.elapsed_time(end))
    plain_eager_ms = statistics.median(eager)
    errs = []
    for m, d, mask in zip(mats, batch.mats, masks):
        ref = dense_attention_fp64_torch(m.q, m.k, m.v, mask, V)
        diff = d.o.double() - ref
        errs.append((float(diff.abs().max()), float(diff.norm() / ref.norm())))
    return {
        "workload": f"attention {samples}x{heads} heads, L~U[512,4096] (x64), d=64, V={V}, s={sparsity}, "
                    "int8 Q/K/V per-tensor 127/max|x|, row-scaled int8 P",
        "matrices": len(mats), "seq_lens": Ls, "nnz_vec": nnz,
        "fused_graph_ms": fused_ms, "plain_graph_ms": plain_graph_ms, "plain_eager_ms": plain_eager_ms,
        "speedup_vs_plain_eager": plain_eager_ms / fused_ms, "speedup_vs_plain_graph": plain_graph_ms / fused_ms,
        "fused_tops": ops / (fused_ms * 1e-3) / 1e12, "algorithmic_bytes": alg,
        "fused_hbm_frac": alg / (fused_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "unfused_bytes": unfused, "unfused_roofline_ms": unfused / HBM_PEAK_GBS / 1e6,
        "plain_graph_hbm_frac": unfused / (plain_graph_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "fused_equals_plain": identical,
        "error_vs_fp64": {"max_abs": max(e[0] for e in errs), "rel_fro_max": max(e[1] for e in errs),
                          "rel_fro_mean": statistics.mean(e[1] for e in errs)},
    }


def _outputs_equal(ref, got):
    return len(ref) == len(got) and all(torch_equal_bits(x, y) for x, y in zip(ref, got))


def torch_equal_bits(x, y):
    import torch
    return x.shape == y.shape and torch.equal(x.contiguous().view(torch.int32), y.contiguous().view(torch.int32))


def serving_trace_bench(n_requests, seed, budget_gib, heads=12, samples=(1, 4), seq=(256, 8192),
                        workload=""):
    """Continuous-batching server vs per-request launches on one request
    trace: saturation throughput (every request queued at t=0) and latency
    under Poisson arrivals at 60% of the batched capacity.  An untimed pass
    per policy keeps every request's outputs, which must be bitwise equal
    across the three policies."""
    import math
    import torch
    from magicube_amd import serving
    budget = int(budget_gib * (1 << 30))
    warm = serving.make_requests(4, math.inf, seed + 1, heads=2, seq=(256, 1024))
    for pol in serving.POLICIES:
        serving.Server(budget, pol).run(warm)
        serving.reset(warm)
    reqs = serving.make_requests(n_requests, math.inf, seed, heads=heads, samples=samples, seq=seq)
    torch.cuda.synchronize()
    out = {"workload": workload, "requests": n_requests, "matrices": sum(len(r.matrices) for r in reqs),
           "hbm_budget_gib": budget_gib}
    # untimed pass per policy with the outputs kept: allocator growth for the
    # per-request arenas and first launches are not serving cost, and the
    # outputs are compared across policies
    ref, identical = None, True
    for pol in serving.POLICIES:
        serving.reset(reqs)
        serving.Server(budget, pol, keep_outputs=True).run(reqs)
        torch.cuda.synchronize()
        got = [[o.clone() for o in r.result()] for r in reqs]
        if ref is None:
            ref = got
        else:
            identical &= all(_outputs_equal(a, b) for a, b in zip(ref, got))
        del got
    del ref
    out["identical"] = bool(identical)
    sat = {}
    for pol in serving.POLICIES:
        serving.reset(reqs)
        sat[pol] = serving.Server(budget, pol).run(reqs)
    out["saturation"] = sat
    out["speedup_req_per_s_vs_plain"] = sat["batched"]["req_per_s"] / sat["plain"]["req_per_s"]
    out["speedup_req_per_s_vs_fused_per_request"] = sat["batched"]["req_per_s"] / sat["fused"]["req_per_s"]
    rate = 0.6 * sat["batched"]["req_per_s"]
    gaps = np.random.default_rng(seed + 2).exponential(1.0 / rate, size=n_requests)
    for r, t in zip(reqs, np.cumsum(gaps) - gaps[0]):
        r.arrival = float(t)
    poi = {}
    for pol in serving.POLICIES:
        serving.reset(reqs)
        poi[pol] = serving.Server(budget, pol).run(reqs)
    out["poisson"] = {"rate_req_per_s": rate, **poi}
    serving.reset(reqs)
    return out


def serving_bench(n_requests=48, seed=202, budget_gib=4.0):
    """The serving trace at its stated size: 12 heads x 1-4 samples,
    L ~ U[256, 8192], s ~ U[0.8, 0.98], int8 or int4 Q/K."""
    return serving_trace_bench(n_requests, seed, budget_gib,
                               workload="attention requests: 12 heads x 1-4 samples, L~U[256,8192] (x64), "
                                        "s~U[0.8,0.98], int8/int4 Q/K, V=8, row-scaled int8 P")


def serving_small_bench(n_requests=192, seed=303, budget_gib=4.0):
    """Small requests, where launch and host work weigh most: 12 heads x 1
    sample, L ~ U[256, 1024]."""
    return serving_trace_bench(n_requests, seed, budget_gib, samples=(1, 1), seq=(256, 1024),
                               workload="attention requests: 12 heads x 1 sample, L~U[256,1024] (x64), "
                                        "s~U[0.8,0.98], int8/int4 Q/K, V=8, row-scaled int8 P")


def time_eager(run, stream, reps=20, warmup=3, rounds=3):
    """Median per-call time (ms) of ``reps`` back-to-back eager calls on
    ``stream``, bracketed by HIP events (library calls are not graph-captured)."""
    import torch
    with torch.cuda.stream(stream):
        for _ in range(warmup):
            run()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    times = []
    for _ in range(rounds):
        with torch.cuda.stream(stream):
            start.record(stream)
            for _ in range(reps):
                run()
            end.record(stream)
        torch.cuda.synchronize()
        times.append(start.elapsed_time(end) / reps)
    return statistics.median(times)


def time_call(run, stream, steps=50):
    """(ms per call, how) for a launch sequence: graph-replayed like the
    headline when the calls can be captured (library calls included), else
    eager calls between HIP events."""
    try:
        ms, _ = time_gpu(run, steps, SECTION_WARMUP, stream)
        return ms, "graph"
    except Exception:
        import torch
        torch.cuda.synchronize()
        return time_eager(run, stream), "eager"


def vendor_bench(seed=7):
    """The headline SpMM and SDDMM beside MI355X's vendor libraries, the roles
    cuSPARSE and cuBLAS play in the reference's baselines
    (baselines/spmm_benchmark.cpp:162-770, baselines/sddmm_benchmark.cpp:33-430).

    Every case is timed the same way: graph-replayed calls on one stream
    (``time_call``; eager calls between HIP events where a library call cannot
    be captured, marked in ``timing``).  ``speedup`` = vendor time / this
    library's time.  Integer vendor results on the same matrix, read after
    the timed replays, are compared with this library's output.
    """
    import torch
    from magicube_amd import sddmm, spmm, synthetic, vendor

    stream = torch.cuda.Stream()
    M = K = 4096
    N, V, s = 512, 8, 0.9
    csr, a, b, a_packed, prep, rhs_np = make_workload(M, K, N, V, s, seed)
    dev = spmm.DeviceSpmm.upload(prep)
    rhs = torch.from_numpy(rhs_np).cuda()
    ours_out = torch.empty((M, N), dtype=torch.int32, device="cuda")

    def ours():
        dev.matmul(rhs, N, out=ours_out, sign=spmm.SIGNED, stream=stream)

    ours_ms, _ = time_call(ours, stream)
    torch.cuda.synchronize()
    ops = 2.0 * csr.nnz_vec * V * N
    cases = []

    def add(name, fmt, dtype, make, out=None, note=None):
        entry = {"library": name, "format": fmt, "dtype": dtype}
        try:
            plan = make()
        except Exception as exc:   # a combination the library does not provide
            entry["error"] = str(exc)
            cases.append(entry)
            return
        ms, how = time_call(plan.run, stream)
        entry.update({"ms": ms, "timing": how, "speedup": ms / ours_ms,
                      "effective_tops": ops / (ms * 1e-3) / 1e12, "workspace_bytes": plan.workspace_bytes})
        if out is not None:
            torch.cuda.synchronize()
            entry["matches_ours"] = bool(torch.equal(out, ours_out))
        if note:
            entry["note"] = note
        plan.close()
        cases.append(entry)

    a8 = torch.from_numpy(a.astype(np.int8)).cuda()
    b8 = torch.from_numpy(b.astype(np.int8).reshape(K, N)).cuda()
    rp, ci, gather = vendor.scalar_csr(csr)
    rp_t, ci_t = torch.from_numpy(rp).cuda(), torch.from_numpy(ci).cuda()
    v8 = a8[torch.from_numpy(gather).cuda()].contiguous()
    c32 = torch.empty((M, N), dtype=torch.int32, device="cuda")
    for alg in vendor.CSR_ALGS:
        add("rocSPARSE", f"CSR ({alg})", "int8", lambda alg=alg: vendor.csr_spmm(
            vendor.I8, alg, M, K, N, rp_t, ci_t, v8, b8, c32, stream), out=c32)
    brp, bci, bv = vendor.bsr_blocks(csr, a.astype(np.int8))
    brp_t, bci_t, bv_t = (torch.from_numpy(x).cuda() for x in (brp, bci, bv))
    add("rocSPARSE", f"BSR {V}x{V} (same matrix, vectors padded)", "int8", lambda: vendor.bsr_spmm(
        vendor.I8, csr.m_vec, K // V, N, V, brp_t, bci_t, bv_t, b8, c32, stream), out=c32)
    urp, uci = vendor.uniform_blocks(csr.m_vec, K, V, 1.0 - s, seed + 5)
    uv = torch.from_numpy(synthetic.random_lanes(uci.size * V * V, 8, True, seed + 6).astype(np.int8)).cuda()
    urp_t, uci_t = torch.from_numpy(urp).cuda(), torch.from_numpy(uci).cuda()
    ublk = f"{uci.size // csr.m_vec} blocks per block row"
    add("rocSPARSE", f"BSR {V}x{V} (uniform blocks, equal density)", "int8", lambda: vendor.bsr_spmm(
        vendor.I8, csr.m_vec, K // V, N, V, urp_t, uci_t, uv, b8, c32, stream),
        note=f"the reference's Blocked-ELL baseline structure, {ublk}")
    dense_a = torch.from_numpy(vendor.dense_lhs(csr, a.astype(np.int8))).cuda()
    b8_t = b8.t().contiguous()
    add("hipBLASLt", "dense GEMM (NN)", "int8", lambda: vendor.gemm(vendor.I8, M, K, N, dense_a, b8, c32,
                                                                  False, stream), out=c32)
    add("hipBLASLt", "dense GEMM (TN)", "int8", lambda: vendor.gemm(vendor.I8, M, K, N, dense_a, b8_t, c32,
                                                                  True, stream), out=c32)
    dense_h = dense_a.to(torch.float16)
    b16, b16_t = b8.to(torch.float16), b8_t.to(torch.float16)
    c16 = torch.empty((M, N), dtype=torch.float16, device="cuda")
    add("hipBLASLt", "dense GEMM (NN)", "fp16", lambda: vendor.gemm(vendor.F16, M, K, N, dense_h, b16, c16,
                                                                  False, stream))
    add("hipBLASLt", "dense GEMM (TN)", "fp16", lambda: vendor.gemm(vendor.F16, M, K, N, dense_h, b16_t, c16,
                                                                  True, stream))
    cf32 = torch.empty((M, N), dtype=torch.float32, device="cuda")
    v16 = v8.to(torch.float16)
    for alg in vendor.CSR_ALGS:
        add("rocSPARSE", f"CSR ({alg})", "fp16", lambda alg=alg: vendor.csr_spmm(
            vendor.F16, alg, M, K, N, rp_t, ci_t, v16, b16, cf32, stream))
    uv16 = uv.to(torch.float16)
    add("rocSPARSE", f"BSR {V}x{V} (uniform blocks, equal density)", "fp16", lambda: vendor.bsr_spmm(
        vendor.F16, csr.m_vec, K // V, N, V, urp_t, uci_t, uv16, b16, cf32, stream),
        note=f"the reference's Blocked-ELL baseline structure, {ublk}")
    spmm_res = {"workload": f"SpMM M=K={M}, N={N}, V={V}, s={s}, signed int8 (L8-R8)", "ours_ms": ours_ms,
                "ours_tops": ops / (ours_ms * 1e-3) / 1e12, "cases": cases}
    del dense_a, dense_h, v8, v16

    # SDDMM: M = N = 4096, K = 64, V = 8, s = 0.9, signed int8
    Ks = 64
    mcsr = synthetic.uniform_block_mask(M, M, V, s, seed + 9)
    mask = sddmm.DeviceMask.upload(mcsr)
    la = synthetic.random_lanes(M * Ks, 8, True, seed + 10)
    lb = synthetic.random_lanes(M * Ks, 8, True, seed + 11)
    lhs = torch.from_numpy(la.astype(np.int8).view(np.int32)).cuda()
    rhs_s = torch.from_numpy(lb.astype(np.int8).view(np.int32)).cuda()
    sres = torch.zeros(mask.output_size(), dtype=torch.int32, device="cuda")

    def ours_s():
        mask.sddmm(8, lhs, rhs_s, Ks, out=sres, stream=stream)

    ours_sms, _ = time_call(ours_s, stream)
    sops = 2.0 * mcsr.nnz_vec * V * Ks
    cases = []
    ops = sops
    ours_ms = ours_sms
    srp, sci, _ = vendor.scalar_csr(mcsr)
    srp_t, sci_t = torch.from_numpy(srp).cuda(), torch.from_numpy(sci).cuda()
    l16 = torch.from_numpy(la.astype(np.float16).reshape(M, Ks)).cuda()
    r16 = torch.from_numpy(lb.astype(np.float16).reshape(M, Ks)).cuda()
    sv = torch.ones(sci.size, dtype=torch.float32, device="cuda")
    add("rocSPARSE", "SDDMM CSR (same mask)", "fp16", lambda: vendor.csr_sddmm(M, M, Ks, srp_t, sci_t, l16, r16,
                                                                             sv, stream))
    l8 = torch.from_numpy(la.astype(np.int8).reshape(M, Ks)).cuda()
    r8_t = torch.from_numpy(lb.astype(np.int8).reshape(M, Ks)).cuda()    # rhs is N x K: the TN form
    dense_c = torch.empty((M, M), dtype=torch.int32, device="cuda")
    add("hipBLASLt", "dense GEMM (TN), full M x N output", "int8",
        lambda: vendor.gemm(vendor.I8, M, Ks, M, l8, r8_t, dense_c, True, stream))
    dense_h16 = torch.empty((M, M), dtype=torch.float16, device="cuda")
    add("hipBLASLt", "dense GEMM (TN), full M x N output", "fp16",
        lambda: vendor.gemm(vendor.F16, M, Ks, M, l16, r16, dense_h16, True, stream))
    sddmm_res = {"workload": f"SDDMM M=N={M}, K={Ks}, V={V}, s={s}, signed int8", "ours_ms": ours_sms,
                 "ours_tops": sops / (ours_sms * 1e-3) / 1e12, "cases": cases}
    return {"spmm": spmm_res, "sddmm": sddmm_res,
            "note": "graph-replayed calls for every entry unless timing says eager; "
                    "speedup = vendor ms / this library's ms"}


def vendor_sweep_bench(seed=61):
    """Speedup over the vendor int8 baselines across sparsity, V and N at the
    reference's DLMC matrix shape (256 vector rows, K = 2304, N in {128, 256},
    unsigned L8-R8), beside the same ratios from the reference's A100 runs
    (Magicube vs cuBLAS int8 dense and cuSPARSE int8 Blocked-ELL).  Every
    entry is timed the same way (``time_call``: graph-replayed calls).  The
    A100 figures time each launch between its own events, launch latency
    included.
    The dense GEMM's time depends on (M, K, N) only; BSR uses uniform V x V
    blocks at the matrix's element density."""
    import torch
    from magicube_amd import spmm, synthetic, vendor
    stream = torch.cuda.Stream()
    K = 2304
    dense_ms = {}
    cases = []
    for N in (128, 256):
        for V in (2, 4, 8):
            M = 256 * V
            a8 = torch.from_numpy(synthetic.random_lanes(M * K, 8, True, seed).astype(np.int8).reshape(M, K)).cuda()
            bt = torch.from_numpy(synthetic.random_lanes(K * N, 8, True, seed + 1).astype(np.int8).reshape(N, K)).cuda()
            c = torch.empty((M, N), dtype=torch.int32, device="cuda")
            plan = vendor.gemm(vendor.I8, M, K, N, a8, bt, c, True, stream)
            dense_ms[(V, N)] = time_call(plan.run, stream)[0]
            plan.close()
            for s in (0.5, 0.7, 0.8, 0.9, 0.95, 0.98):
                csr, a, b, _, prep, rhs_np = make_workload(M, K, N, V, s, seed + 2, 8, 8, signed=False)
                dev = spmm.DeviceSpmm.upload(prep)
                rhs = torch.from_numpy(rhs_np).cuda()
                out = torch.empty((M, N), dtype=torch.int32, device="cuda")

                def run():
                    dev.matmul(rhs, N, out=out, sign=spmm.UNSIGNED, stream=stream)

                with torch.cuda.stream(stream):
                    run()
                torch.cuda.synchronize()
                ok = checksum_ok(out.cpu().numpy(), csr, a, b, K, N, V)
                ours = time_call(run, stream)[0]
                urp, uci = vendor.uniform_blocks(csr.m_vec, K, V, 1.0 - s, seed + 3)
                uv = torch.from_numpy(synthetic.random_lanes(uci.size * V * V, 8, True, seed + 4).astype(np.int8)).cuda()
                b8 = torch.from_numpy(b.astype(np.uint8).view(np.int8).reshape(K, N)).cuda()
                bsr = vendor.bsr_spmm(vendor.I8, csr.m_vec, K // V, N, V, torch.from_numpy(urp).cuda(),
                                      torch.from_numpy(uci).cuda(), uv, b8, out, stream)
                bsr_ms = time_call(bsr.run, stream)[0]
                bsr.close()
                pub = PUBLISHED_DLMC_A100.get((s, V, N))
                cases.append({
                    "sparsity": s, "V": V, "N": N, "nnz_vec": csr.nnz_vec, "verified": ok,
                    "ours_ms": ours, "hipblaslt_int8_ms": dense_ms[(V, N)], "rocsparse_bsr_int8_ms": bsr_ms,
                    "speedup_vs_dense": dense_ms[(V, N)] / ours, "speedup_vs_bsr": bsr_ms / ours,
                    "a100_magicube_ms": pub[0] if pub else None,
                    "a100_speedup_vs_cublas_int8": pub[1] / pub[0] if pub else None,
                    "a100_speedup_vs_cusparse_int8": pub[2] / pub[0] if pub else None,
                    "speedup_vs_a100_magicube": pub[0] / ours if pub else None})
    return {"workload": "SpMM L8-R8 unsigned, m_vec=256, K=2304, N in {128, 256}, V in {2, 4, 8}, "
                        "s in {0.5 .. 0.98} (Poisson mask)",
            "a100_source": "plot/examples/spmm_magicube_8b8b.txt, spmm_cublas_int8.txt, spmm_cusparse_int8.txt",
            "cases": cases}


def load_traffic(name):
    """HBM bytes per launch recorded in a profiles/ summary ("hbm_bytes_per_launch: N")."""
    path = os.path.join(HERE, "profiles", name)
    try:
        with open(path) as f:
            for line in f:
                if line.strip().startswith("hbm_bytes_per_launch:"):
                    return int(line.split(":", 1)[1].strip())
    except (OSError, ValueError):
        pass
    return None


def calibrate_copy(mib=64, iters=20):
    """A plain device-to-device copy of known size, for scaling the
    FETCH_SIZE / WRITE_SIZE counters of a profiled run."""
    import torch
    n = mib * 1024 * 1024 // 4
    x = torch.ones(n, dtype=torch.int32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(iters):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"calibration_copy_bytes": n * 4, "iters": iters}))


def _secondary(sections):
    """A few figures of the other sections for the one-line summary."""
    sec = {}
    rs = sections.get("reference_shape")
    if rs:
        sec["spmm_ref_shape_2048x2304x512"] = {"tops": rs["tops"], "x_a100": rs["ratio_vs_a100"]}
    sd = sections.get("sddmm")
    if sd:
        for c in sd["cases"]:
            if c["bits"] == 8 and c["K"] == 64 and c["sparsity"] == 0.9:
                sec["sddmm_i8_k64_s090"] = {"us": c["us"], "tops": c["tops"], "frac": c["hbm_frac"],
                                            "verified": c["verified"]}
            if c["bits"] == 4 and c["K"] == 128 and c["sparsity"] == 0.95:
                sec["sddmm_i4_k128_s095"] = {"us": c["us"], "tops": c["tops"], "frac": c["hbm_frac"],
                                             "verified": c["verified"]}
    at = sections.get("attention")
    if at:
        sec["attention_8x16"] = {"fused_ms": at["fused_graph_ms"], "x_plain_eager": at["speedup_vs_plain_eager"],
                                 "tops": at["fused_tops"], "fused_equals_plain": at["fused_equals_plain"],
                                 "rel_fro_fp64_max": at["error_vs_fp64"]["rel_fro_max"]}
    sv = sections.get("serving")
    if sv:
        for name in ("serving", "serving_small"):
            t = sections.get(name)
            if not t:
                continue
            sat, poi = t["saturation"], t["poisson"]
            pols = list(sat)
            # lists in the order of "policies": saturation req/s, Poisson p50/p99
            sec[name] = {"policies": ",".join(pols), "req_s": [sat[p]["req_per_s"] for p in pols],
                         "p50_ms": [poi[p]["p50_ms"] for p in pols], "p99_ms": [poi[p]["p99_ms"] for p in pols],
                         "identical": t.get("identical")}
    en = sections.get("encoder")
    if en:
        for c in en["cases"]:
            if c.get("batch") == 8 and c.get("lhs_pre") == 8 and c.get("rhs_pre") == 8:
                sec["encoder_b8_l8r8_ms"] = c["eager_ms"]
    return sec


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5000,
                    help="untimed steps before the headline timing (also brings the GPU to its working clocks)")
    ap.add_argument("--warmup-s", type=float, default=0.3,
                    help="minimum seconds of back-to-back untimed replays before the headline timing, "
                         "whatever --warmup is")
    ap.add_argument("--sections-out", default=os.path.join(tempfile.gettempdir(), "magicube_bench_sections.json"),
                    help="file for the full per-section results (stdout carries one short line)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-reference-shape", action="store_true")
    ap.add_argument("--no-attention", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-sddmm", action="store_true")
    ap.add_argument("--no-encoder", action="store_true", help="skip the 4-layer encoder benchmark")
    ap.add_argument("--no-vendor", action="store_true", help="skip the rocSPARSE / hipBLASLt comparison")
    ap.add_argument("--no-ablation", action="store_true", help="skip the SpMM ablation")
    ap.add_argument("--no-precision", action="store_true",
                    help="skip the every-pair sweep and the published-shape SDDMM")
    ap.add_argument("--headline-only", action="store_true", help="only the headline SpMM (for profiling)")
    ap.add_argument("--config", type=int, default=None, help="explicit SpMM tile config index")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 64 MiB device copy (counter calibration for rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the SpMM would shard by vector-row blocks with B replicated and no exchange)", file=sys.stderr)
        sys.exit(2)
    if args.headline_only:
        for k in ("cpu_baseline", "reference_shape", "attention", "serving", "sddmm", "encoder", "vendor",
                  "ablation", "precision"):
            setattr(args, "no_" + k, True)

    import torch
    from magicube_amd import spmm

    M = K = 4096
    N, V, sparsity, seed = 512, 8, 0.9, 7
    csr, a, b, a_packed, prep, rhs_np = make_workload(M, K, N, V, sparsity, seed)
    dev = spmm.DeviceSpmm.upload(prep)
    rhs = torch.from_numpy(rhs_np).cuda()
    out = torch.empty((M, N), dtype=torch.int32, device="cuda")
    stream = torch.cuda.Stream()

    def run_once():
        dev.matmul(rhs, N, out=out, sign=spmm.SIGNED, config=args.config, stream=stream)

    with torch.cuda.stream(stream):
        run_once()
    torch.cuda.synchronize()
    verified = checksum_ok(out.cpu().numpy(), csr, a, b, K, N, V)
    t, _ = time_gpu(run_once, args.steps, args.warmup, stream, warm_s=args.warmup_s, detail=True)
    ms = t["ms"]
    nnz = csr.nnz_vec
    tops = 2.0 * nnz * V * N / (ms * 1e-3) / 1e12
    alg = algorithmic_bytes(nnz, csr.m_vec, M, K, N, V)
    achieved = alg / (ms * 1e-3) / 1e9
    traffic = load_traffic("spmm_l8r8_m4096_n512_v8_s090_pmc.md")
    # the result after the timed replays is still the exact product
    torch.cuda.synchronize()
    verified = verified and checksum_ok(out.cpu().numpy(), csr, a, b, K, N, V)

    result = {
        "metric": "SpMM/SDDMM int8 TOPS & roofline %; sparse-attn req/s with p50/p99 latency",
        "value": tops,
        "unit": "TOPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8",
        "data": "synthetic: seeded Poisson vector-sparse mask, full-range signed int8",
        "config": {"workload": "spmm_l8r8 M=K=4096 N=512 V=8 s=0.9", "signed": True,
                   "seed": seed, "nnz_vec": nnz, "tile_config": args.config},
        "verified": verified,
        "timing": {"replays": 5, "first_ms": t["first_ms"], "host_wall_ms": t["host_wall_ms"],
                   "warmup_steps_run": t["warmup_steps"], "warmup_s": t["warmup_s"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes": alg, "mfma_useful_frac": tops / MFMA_INT8_PEAK_TOPS},
    }
    sections = {}

    def section(name, fn):
        # each section's progress on stderr, so a long run shows it is alive
        t0 = time.perf_counter()
        print(f"bench: {name} ...", file=sys.stderr, flush=True)
        sections[name] = fn()
        print(f"bench: {name} done in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)

    if not args.no_reference_shape:
        def ref_shape():
            # the published A100 shape: 256 vector rows (M = 2048), K = 2304, N = 512, V = 8, s = 0.9
            rc, ra, rb, rap, rprep, rrhs = make_workload(2048, 2304, 512, 8, 0.9, seed=11)
            rdev = spmm.DeviceSpmm.upload(rprep)
            rr = torch.from_numpy(rrhs).cuda()
            ro = torch.empty((2048, 512), dtype=torch.int32, device="cuda")

            def run_ref():
                rdev.matmul(rr, 512, out=ro, sign=spmm.SIGNED, stream=stream)

            rms, _ = time_gpu(run_ref, args.steps, SECTION_WARMUP, stream)
            rtops = 2.0 * rc.nnz_vec * 8 * 512 / (rms * 1e-3) / 1e12
            return {"M": 2048, "K": 2304, "N": 512, "V": 8, "sparsity": 0.9, "ms_per_step": rms, "tops": rtops,
                    "a100_published_tops": REFERENCE_A100_TOPS, "ratio_vs_a100": rtops / REFERENCE_A100_TOPS,
                    "note": "synthetic pattern at the published DLMC shape"}
        section("reference_shape", ref_shape)

    if not args.no_sddmm:
        section("sddmm", lambda: sddmm_bench(max(1, args.steps // 2), SECTION_WARMUP))

    if not args.no_precision:
        section("precision", lambda: precision_bench(max(1, args.steps // 2), SECTION_WARMUP))
        if not args.no_sddmm:
            section("sddmm_published_shape", lambda: sddmm_published_bench(max(1, args.steps // 2), SECTION_WARMUP))

    if not args.no_attention:
        section("attention", lambda: attention_bench(max(1, args.steps // 10), SECTION_WARMUP // 2))

    if not args.no_serving:
        section("serving", serving_bench)
        section("serving_small", serving_small_bench)

    if not args.no_encoder:
        section("encoder", encoder_bench)

    if not args.no_vendor:
        section("vendor", vendor_bench)
        section("vendor_sweep", vendor_sweep_bench)

    if not args.no_ablation:
        section("ablation", lambda: ablation_bench(max(1, args.steps // 4), SECTION_WARMUP // 2))

    if not args.no_cpu_baseline:
        section("cpu_baseline", lambda: cpu_baseline(csr, a_packed, rhs_np, M, K, N, V))
        result["cpu_baseline"] = sections["cpu_baseline"]

    result["secondary"] = _secondary(sections)
    if sections:
        try:
            with open(args.sections_out, "w") as f:
                json.dump({"headline": result, **sections}, f)
            result["sections_file"] = args.sections_out
        except OSError as e:
            result["sections_file"] = f"not written: {e}"
    line = json.dumps(_rounded(result), separators=(",", ":"))
    # the line stays under 2 KB: secondary figures that do not fit are left
    # to the sections file, least central first
    for key in ("encoder_b8_l8r8_ms", "serving", "sddmm_i4_k128_s095", "serving_small", "attention_8x16"):
        if len(line) <= 2000:
            break
        result.get("secondary", {}).pop(key, None)
        line = json.dumps(_rounded(result), separators=(",", ":"))
    print(line, flush=True)


def _rounded(x, digits=4):
    """Floats to ``digits`` significant digits, so the summary line stays short."""
    if isinstance(x, float):
        return float(f"{x:.{digits}g}")
    if isinstance(x, dict):
        return {k: _rounded(v, digits) for k, v in x.items()}
    if isinstance(x, (list, tuple)):
        return [_rounded(v, digits) for v in x]
    return x


if __name__ == "__main__":
    main()
