"""Pins the CPU oracle (oracle/oracle.c) before anything is checked against it.

* the generator restatement against libstdc++'s <random> (g++ here);
* the plain SpMM/SDDMM sums against an independent numpy dense product;
* the reference-literal host checks (compute_ref_integers,
  Host_sddmm_integers) against the plain sums on the reference's own
  generator data, including the top-lane sign-extension quirk of the literal
  B extraction (SpMM/SpMM/spmm_benchmark.cpp:106);
* the published-log invariants (plot/examples/spmm_abl_study.txt:10-14,23-26);
* planted mistakes that each of these comparisons must catch.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from magicube_amd import synthetic

# (preA, preB, V) -> (TypeA bytes, preA container, preA_cut, scaleA): the
# dispatch table of main(), SpMM/SpMM/spmm_benchmark.cpp:809-831
REF_TABLE = {
    (4, 4, 8): (4, 4, 4, 1), (4, 4, 4): (2, 4, 4, 1), (4, 4, 2): (1, 4, 4, 1),
    (8, 4, 8): (8, 8, 8, 1), (8, 4, 4): (4, 8, 8, 1), (8, 4, 2): (2, 8, 8, 1),
    (12, 4, 8): (8, 16, 12, 2), (12, 4, 4): (8, 16, 12, 1), (12, 4, 2): (4, 16, 12, 1),
    (16, 4, 8): (8, 16, 16, 2), (16, 4, 4): (8, 16, 16, 1), (16, 4, 2): (4, 16, 16, 1),
    (8, 8, 8): (8, 8, 8, 1), (8, 8, 4): (4, 8, 8, 1), (8, 8, 2): (2, 8, 8, 1),
    (12, 8, 8): (8, 16, 12, 2), (12, 8, 4): (8, 16, 12, 1), (12, 8, 2): (4, 16, 12, 1),
    (16, 8, 8): (8, 16, 16, 2), (16, 8, 4): (8, 16, 16, 1), (16, 8, 2): (4, 16, 16, 1),
    (16, 16, 8): (2, 16, 16, 8), (16, 16, 4): (2, 16, 16, 4), (16, 16, 2): (2, 16, 16, 2),
}


def dense_spmm(m_vec, V, K, N, ro, ci, a, b):
    """Independent restatement: expand A to dense int64 and multiply (mod 2^32)."""
    A = np.zeros((m_vec * V, K), dtype=np.int64)
    a = np.asarray(a, dtype=np.int64).reshape(-1, V)
    for i in range(m_vec):
        for j in range(ro[i], ro[i + 1]):
            A[i * V:(i + 1) * V, ci[j]] += a[j]
    C = A @ np.asarray(b, dtype=np.int64).reshape(K, N)
    return (C & 0xFFFFFFFF).astype(np.uint32).view(np.int32)


def small_mask(m_vec, K, V, density, seed, empty_rows=True):
    csr = synthetic.poisson_block_mask(m_vec * V, K, V, 1.0 - density, seed)
    if empty_rows and m_vec > 3:
        # force a few empty rows (edge case the reference handles by storing zeros)
        ro = csr.row_offsets.copy()
        counts = np.diff(ro)
        counts[[0, m_vec // 2]] = 0
        cols = [csr.col_indices[ro[i]:ro[i + 1]] for i in range(m_vec)]
        cols[0] = cols[0][:0]
        cols[m_vec // 2] = cols[m_vec // 2][:0]
        ro = np.concatenate([[0], np.cumsum([c.size for c in cols])]).astype(np.int32)
        csr.row_offsets = ro
        csr.col_indices = np.concatenate(cols).astype(np.int32)
    return csr


# --------------------------------------------------------------------------
# generator vs libstdc++
# --------------------------------------------------------------------------

_CPP_GEN = r"""
#include <cstdio>
#include <random>
int main(int argc, char** argv) {
    int n = 257;
    {   // raw canonical floats of a fresh engine
        std::default_random_engine e;
        std::uniform_real_distribution<float> u(0.0, 1.0);
        for (int i = 0; i < n; ++i) { float x = u(e); unsigned b; __builtin_memcpy(&b, &x, 4); std::printf("f %u\n", b); }
    }
    {   // int elements: one discarded draw, then 0x11111111 + int(117*u)
        std::default_random_engine e;
        std::uniform_real_distribution<float> u(0.0, 1.0);
        for (int i = 0; i < n; ++i) { volatile float t = 2147483647.0 * u(e); (void)t;
            int v = 0x11111111 + int(117.0 * u(e)); std::printf("i %d\n", v); }
    }
    {   // short elements: short(float(10240*u))
        std::default_random_engine e;
        std::uniform_real_distribution<float> u(0.0, 1.0);
        for (int i = 0; i < n; ++i) { float t = 10240.0 * u(e); std::printf("s %d\n", (int)(short)t); }
    }
    {   // char elements: char(float(127*u))
        std::default_random_engine e;
        std::uniform_real_distribution<float> u(0.0, 1.0);
        for (int i = 0; i < n; ++i) { float t = 127.0 * u(e); std::printf("c %d\n", (int)(signed char)t); }
    }
    return 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_generator_matches_libstdcxx(orc, tmp_path):
    src = tmp_path / "gen.cpp"
    src.write_text(_CPP_GEN)
    exe = tmp_path / "gen"
    subprocess.run(["g++", "-O2", "-o", str(exe), str(src)], check=True)
    lines = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {"f": [], "i": [], "s": [], "c": []}
    for ln in lines:
        if ln:
            k, v = ln.split()
            got[k].append(int(v))
    n = 257
    f = np.array(got["f"], dtype=np.uint32).view(np.float32)
    assert np.array_equal(orc.canonical_stream(n).view(np.uint32), f.view(np.uint32))
    assert np.array_equal(orc.make_dense_matrix(4, n), np.array(got["i"], dtype=np.int32))
    assert np.array_equal(orc.make_dense_matrix(2, n), np.array(got["s"], dtype=np.int16))
    assert np.array_equal(orc.make_dense_matrix(1, n), np.array(got["c"], dtype=np.int8))


def test_generator_value_ranges(orc):
    # bm_test_utils.h:136-156: every int word is 0x11111111 + [0,116]
    w = orc.make_dense_matrix(4, 10000)
    assert w.min() >= 0x11111111 and w.max() <= 0x11111111 + 116
    ll = orc.make_dense_matrix(8, 1000)
    assert (ll >= 0x1111111111111111).all() and (ll <= 0x1111111111111111 + 116).all()
    s = orc.make_dense_matrix(2, 1000)
    assert s.min() >= 0 and s.max() < 10240
    # the SDDMM copy draws long long from the full positive range
    lls = orc.make_dense_matrix(8, 1000, sddmm=True)
    assert (lls >= 0).all() and lls.max() > (1 << 60)
    # first canonical draw: minstd_rand0 from seed 1 gives 16807
    assert orc.canonical_stream(1)[0] == np.float32(16806) / np.float32(2 ** 31)


# --------------------------------------------------------------------------
# plain sums vs independent dense product
# --------------------------------------------------------------------------

@pytest.mark.parametrize("V", [2, 4, 8])
@pytest.mark.parametrize("bits,signed", [(8, True), (8, False), (4, True), (16, True), (16, False)])
def test_spmm_plain_matches_dense(orc, V, bits, signed):
    m_vec, K, N = 12, 48, 24
    csr = small_mask(m_vec, K, V, 0.3, seed=V * 100 + bits)
    nnz = csr.nnz_vec
    a = synthetic.random_lanes(nnz * V, bits, signed, seed=1)
    b = synthetic.random_lanes(K * N, 8, True, seed=2)
    got = orc.spmm_plain(m_vec, V, N, csr.row_offsets, csr.col_indices, a, b)
    exp = dense_spmm(m_vec, V, K, N, csr.row_offsets, csr.col_indices, a, b)
    assert np.array_equal(got, exp)


def test_spmm_plain_wraps_mod_2_32(orc):
    # 3 products of 2^30 overflow int32; the oracle keeps the sum modulo 2^32
    ro = np.array([0, 3], np.int32)
    ci = np.array([0, 1, 2], np.int32)
    a = np.full(3 * 2, 1 << 15, np.int32)
    b = np.full(3 * 4, 1 << 15, np.int32)
    got = orc.spmm_plain(1, 2, 4, ro, ci, a, b)
    assert (got == np.int32(np.uint32((3 << 30) & 0xFFFFFFFF).view(np.int32))).all()


def test_spmm_special_cases(orc):
    m_vec, K, N, V = 6, 16, 16, 4
    csr = small_mask(m_vec, K, V, 0.5, seed=3)
    a = synthetic.random_lanes(csr.nnz_vec * V, 8, True, seed=4)
    # one-hot B (identity): C row block = scatter of A's vectors into their columns
    eye = np.eye(K, N, dtype=np.int32)
    C = orc.spmm_plain(m_vec, V, N, csr.row_offsets, csr.col_indices, a, eye)
    for i in range(m_vec):
        for j in range(csr.row_offsets[i], csr.row_offsets[i + 1]):
            assert np.array_equal(C[i * V:(i + 1) * V, csr.col_indices[j]], a[j * V:(j + 1) * V])
    # all-ones B: every column holds the row sums of A
    ones = np.ones(K * N, np.int32)
    C = orc.spmm_plain(m_vec, V, N, csr.row_offsets, csr.col_indices, a, ones)
    av = a.reshape(-1, V)
    for i in range(m_vec):
        s = av[csr.row_offsets[i]:csr.row_offsets[i + 1]].sum(axis=0) if csr.row_offsets[i + 1] > csr.row_offsets[i] \
            else np.zeros(V, np.int64)
        assert (C[i * V:(i + 1) * V] == s[:, None]).all()
    # empty rows (forced by small_mask) give zeros
    assert (C[0:V] == 0).all()


@pytest.mark.parametrize("V", [2, 8])
def test_sddmm_plain_matches_dense(orc, V):
    m_vec, N, K = 10, 40, 32
    csr = small_mask(m_vec, N, V, 0.3, seed=7)
    pairs, an = orc.align_row_offsets(csr.row_offsets, 8)
    cols = orc.align_col_indices(csr.row_offsets, csr.col_indices, pairs, an)
    lhs = synthetic.random_lanes(m_vec * V * K, 8, True, seed=8)
    rhs = synthetic.random_lanes(N * K, 8, True, seed=9)
    out = orc.sddmm_plain(m_vec, V, K, pairs, cols, lhs, rhs, an * V, 8)
    D = lhs.reshape(-1, K).astype(np.int64) @ rhs.reshape(N, K).astype(np.int64).T
    for i in range(m_vec):
        for j in range(pairs[2 * i], pairs[2 * i + 1]):
            for v in range(V):
                assert out[(j // 8) * 8 * V + 8 * v + j % 8] == np.int32(D[i * V + v, cols[j]])
        for j in range(pairs[2 * i + 1], pairs[2 * i] + (-(pairs[2 * i] - pairs[2 * i + 1]) + 7) // 8 * 8):
            for v in range(V):
                assert out[(j // 8) * 8 * V + 8 * v + j % 8] == 0


# --------------------------------------------------------------------------
# reference-literal checks vs plain sums on reference data
# --------------------------------------------------------------------------

def _ref_case(orc, preA, preB, V, m_vec=16, K=64, N=32, seed=11):
    tb, pa, cut, scale = REF_TABLE[(preA, preB, V)]
    csr = small_mask(m_vec, K, V, 0.25, seed)
    nnz = csr.nnz_vec
    count = nnz * V * scale * pa // (tb * 8)
    vals = orc.make_dense_matrix(tb, count)
    B = orc.make_dense_matrix(4, K * N * preB // 32)
    return csr, vals, B, (tb, pa, cut, scale)


@pytest.mark.parametrize("key", sorted(REF_TABLE))
def test_ref_check_equals_plain_unsigned_on_reference_data(orc, key):
    preA, preB, V = key
    m_vec, K, N = 16, 64, 32
    csr, vals, B, (tb, pa, cut, scale) = _ref_case(orc, preA, preB, V, m_vec, K, N)
    C_ref, flops = orc.spmm_ref_check(vals, tb, B, m_vec * V, K, N, pa, cut, preB, V, csr.row_offsets,
                                      csr.col_indices, m_vec, scale)
    nnz = csr.nnz_vec
    used = vals.view(np.uint8)[: nnz * V * pa // 8]
    a = orc.unpack_lanes(used, pa, nnz * V, signed=False) & ((1 << cut) - 1)
    b = orc.unpack_lanes(B.view(np.uint8), preB, K * N, signed=False)
    assert np.array_equal(C_ref, orc.spmm_plain(m_vec, V, N, csr.row_offsets, csr.col_indices, a, b))
    assert flops == 2.0 * nnz * V * N


def test_ref_check_top_lane_quirk(orc):
    # Full-range B: the literal extraction sign-extends the top lane of each
    # int32 (int arithmetic, spmm_benchmark.cpp:106) and zero-extends A.
    m_vec, K, N, V = 8, 32, 16, 4
    csr = small_mask(m_vec, K, V, 0.4, seed=5)
    nnz = csr.nnz_vec
    a = synthetic.random_lanes(nnz * V, 8, False, seed=6)
    b = synthetic.random_lanes(K * N, 8, False, seed=7)
    A_words = orc.pack_lanes(a, 8).view(np.int32)
    B_words = orc.pack_lanes(b, 8).view(np.int32)
    C_ref, _ = orc.spmm_ref_check(A_words, 4, B_words, m_vec * V, K, N, 8, 8, 8, V, csr.row_offsets,
                                  csr.col_indices, m_vec, 1)
    b_quirk = b.copy().reshape(-1, 4)
    b_quirk[:, 3] = np.where(b_quirk[:, 3] >= 128, b_quirk[:, 3] - 256, b_quirk[:, 3])
    assert np.array_equal(C_ref, orc.spmm_plain(m_vec, V, N, csr.row_offsets, csr.col_indices, a, b_quirk.ravel()))
    # ...and therefore differs from the all-unsigned reading on such data
    assert not np.array_equal(C_ref, orc.spmm_plain(m_vec, V, N, csr.row_offsets, csr.col_indices, a, b))


# (preA, preB, V) entries of the dispatch table with one container per vector
# and no cut: the cases the row-parallel loop handles
OMP_KEYS = sorted(k for k, (tb, pa, cut, scale) in REF_TABLE.items() if scale == 1 and cut == pa == k[0])


@pytest.mark.parametrize("data", ["reference", "full_range"])
@pytest.mark.parametrize("key", OMP_KEYS)
def test_ref_check_omp_equals_single_thread(orc, key, data):
    """The OpenMP loop bench.py times as the parallel CPU baseline gives the
    literal single-thread check's C and flop count bit for bit, on the
    reference generator's data and on full-range lanes (where the top B lane
    of each int32 sign-extends, spmm_benchmark.cpp:106)."""
    preA, preB, V = key
    m_vec, K, N = 24, 96, 64
    tb = REF_TABLE[key][0]
    if data == "reference":
        csr, vals, B, _ = _ref_case(orc, preA, preB, V, m_vec, K, N, seed=41)
    else:
        csr = small_mask(m_vec, K, V, 0.3, seed=43)
        count = csr.nnz_vec * V * preA // (tb * 8)
        rng = np.random.default_rng(44)
        vals = rng.integers(0, 256, size=count * tb, dtype=np.uint8).view({1: np.int8, 2: np.int16, 4: np.int32,
                                                                           8: np.int64}[tb])
        B = rng.integers(0, 256, size=K * N * preB // 8, dtype=np.uint8).view(np.int32)
    C1, f1 = orc.spmm_ref_check(vals, tb, B, m_vec * V, K, N, preA, preA, preB, V, csr.row_offsets,
                                csr.col_indices, m_vec, 1)
    for threads in (1, 3, 8):
        Cp, fp = orc.spmm_ref_check_omp(vals, tb, B, m_vec * V, K, N, preA, preB, V, csr.row_offsets,
                                        csr.col_indices, m_vec, threads)
        assert np.array_equal(Cp, C1), (key, data, threads)
        assert fp == f1
    if data == "full_range":
        # the data reaches the quirk: the all-unsigned product differs
        a = orc.unpack_lanes(vals.view(np.uint8), preA, csr.nnz_vec * V, signed=False)
        b = orc.unpack_lanes(B.view(np.uint8), preB, K * N, signed=False)
        assert not np.array_equal(C1, orc.spmm_plain(m_vec, V, N, csr.row_offsets, csr.col_indices, a, b))


def test_ref_check_omp_planted_mistake_is_caught(orc):
    """A column index off by one in the parallel loop's input changes its C,
    so the comparison above would catch such a slip."""
    key = (8, 8, 8)
    csr, vals, B, _ = _ref_case(orc, 8, 8, 8, 16, 64, 32, seed=45)
    C1, _ = orc.spmm_ref_check(vals, 8, B, 128, 64, 32, 8, 8, 8, 8, csr.row_offsets, csr.col_indices, 16, 1)
    bad_ci = (csr.col_indices + 1) % 64
    Cp, _ = orc.spmm_ref_check_omp(vals, 8, B, 128, 64, 32, 8, 8, 8, csr.row_offsets, bad_ci, 16, 4)
    assert not np.array_equal(Cp, C1), key


@pytest.mark.parametrize("V", [2, 8])
def test_published_counter_invariant(orc, V):
    # plot/examples/spmm_abl_study.txt:10-14 (V=2) and :22-26 (V=8), L16-R8:
    # "total Gflops" = 2*nnz*V*N and counter = M*N (every output > 0).
    m_vec, K, N = 16, 96, 64
    csr, vals, B, (tb, pa, cut, scale) = _ref_case(orc, 16, 8, V, m_vec, K, N, seed=21)
    # make every row non-empty, as in the DLMC matrices of the log
    assert True
    C_ref, flops = orc.spmm_ref_check(vals, tb, B, m_vec * V, K, N, pa, cut, 8, V, csr.row_offsets,
                                      csr.col_indices, m_vec, scale)
    nonempty = np.repeat(np.diff(csr.row_offsets) > 0, V)
    assert (C_ref[nonempty] > 0).all()
    assert flops == 2.0 * csr.nnz_vec * V * N
    # the log's own figures: nnz 176947, V=2, N=512 -> 0.362387 Gflop; V=8 -> 1.44955
    assert round(2 * 176947 * 2 * 512 / 1e9, 6) == 0.362387
    assert round(2 * 176947 * 8 * 512 / 1e9, 5) == 1.44955


@pytest.mark.parametrize("preA,V", [(8, 8), (8, 2), (4, 8), (16, 8), (16, 2)])
def test_sddmm_ref_check_equals_plain_unsigned(orc, preA, V):
    m_vec, N, K = 12, 48, 64
    csr = small_mask(m_vec, N, V, 0.3, seed=13)
    pairs, an = orc.align_row_offsets(csr.row_offsets, 8)
    cols = orc.align_col_indices(csr.row_offsets, csr.col_indices, pairs, an)
    lhs = orc.make_dense_matrix(4, m_vec * V * K * preA // 32)
    rhs = orc.make_dense_matrix(4, N * K * preA // 32)
    out_ref, _ = orc.sddmm_ref_check(lhs, rhs, m_vec * V, K, N, preA, preA, V, pairs, cols, m_vec, 8, an * V)
    l = orc.unpack_lanes(lhs.view(np.uint8), preA, m_vec * V * K, signed=False)
    r = orc.unpack_lanes(rhs.view(np.uint8), preA, N * K, signed=False)
    out_plain = orc.sddmm_plain(m_vec, V, K, pairs, cols, l, r, an * V, 8)
    assert np.array_equal(out_ref, out_plain)


# --------------------------------------------------------------------------
# preparation layouts: alignment, shuffle, value image
# --------------------------------------------------------------------------

def test_align_and_shuffle(orc):
    csr = small_mask(20, 100, 4, 0.3, seed=17)
    for align in (8, 16, 32):
        pairs, an = orc.align_row_offsets(csr.row_offsets, align)
        assert an % align == 0
        starts, ends = pairs[0::2], pairs[1::2]
        assert (starts % align == 0).all()
        assert np.array_equal(ends - starts, np.diff(csr.row_offsets))
        cols = orc.align_col_indices(csr.row_offsets, csr.col_indices, pairs, an)
        assert (cols == -1).sum() == an - csr.nnz_vec
        sh = orc.shuffle_col_indices(cols)
        g = sh.reshape(-1, 8)
        assert np.array_equal(g[:, :4], cols.reshape(-1, 8)[:, 0::2])
        assert np.array_equal(g[:, 4:], cols.reshape(-1, 8)[:, 1::2])


def read_prepared(prep, lhs_bits, rhs_bits, V, pairs, aligned_num):
    """Decodes the prepared value image by the documented per-group layout."""
    mma_k = 32 if (lhs_bits == 4 or rhs_bits == 4) else 16
    out = np.zeros((aligned_num, V), dtype=np.int64)
    b = prep.astype(np.int64)
    cb = 16 if lhs_bits == 12 else lhs_bits          # 12-bit lanes use 16-bit containers
    gb = mma_k * V * cb // 8
    for s in range(aligned_num):
        G, u = divmod(s, mma_k)
        base = G * gb
        for v in range(V):
            if lhs_bits == 8 and rhs_bits == 8:
                out[s, v] = b[base + 16 * v + u]
            elif lhs_bits >= 12 and rhs_bits == 4:
                # nibble planes q (bits 4q..4q+3) at 16*V*q
                out[s, v] = sum(((b[base + 16 * V * q + 16 * v + u // 2] >> (4 * (u & 1))) & 0xF) << (4 * q)
                                for q in range(lhs_bits // 4))
            elif lhs_bits >= 12:
                out[s, v] = b[base + 16 * v + u] | (b[base + 16 * V + 16 * v + u] << 8)
            elif lhs_bits == 8:
                lo = (b[base + 16 * v + u // 2] >> (4 * (u & 1))) & 0xF
                hi = (b[base + 16 * V + 16 * v + u // 2] >> (4 * (u & 1))) & 0xF
                out[s, v] = lo | (hi << 4)
            else:
                out[s, v] = (b[base + 16 * v + u // 2] >> (4 * (u & 1))) & 0xF
    return out


@pytest.mark.parametrize("V", [2, 4, 8])
@pytest.mark.parametrize("preA,preB", [(8, 8), (16, 8), (8, 4), (4, 4), (12, 8), (16, 16), (12, 4), (16, 4)])
def test_prepared_image_encodes_values(orc, V, preA, preB):
    m_vec, K = 10, 80
    tb, pa, cut, scale = REF_TABLE[(preA, preB, V)]
    csr = small_mask(m_vec, K, V, 0.3, seed=V + preA)
    nnz = csr.nnz_vec
    lanes = synthetic.random_lanes(nnz * V, preA, False, seed=3)
    if preA == 12:
        # 16-bit containers with junk in bits 12..15, which the layout drops
        junk = synthetic.random_lanes(nnz * V, 4, False, seed=4)
        vals = orc.pack_lanes(lanes | (junk << 12), 16)
    else:
        vals = orc.pack_lanes(lanes, preA)
    vals = np.concatenate([vals, np.zeros((-vals.size) % tb, np.uint8)]).view({1: np.int8, 2: np.int16,
                                                                                 4: np.int32, 8: np.int64}[tb])
    mma_k = 32 if (preA == 4 or preB == 4) else 16
    pairs, an = orc.align_row_offsets(csr.row_offsets, mma_k)
    prep = orc.prepare_values(vals, tb, scale, csr.row_offsets, pairs, an, V, pa, cut, preB)
    dec = read_prepared(prep, preA, preB, V, pairs, an)
    lv = lanes.reshape(-1, V)
    for i in range(m_vec):
        s0, e0 = pairs[2 * i], pairs[2 * i + 1]
        assert np.array_equal(dec[s0:e0], lv[csr.row_offsets[i]:csr.row_offsets[i + 1]])
        pad_end = s0 + -(-(e0 - s0) // mma_k) * mma_k
        assert (dec[e0:pad_end] == 0).all()


# --------------------------------------------------------------------------
# planted mistakes: each comparison above must notice them
# --------------------------------------------------------------------------

@pytest.mark.parametrize("mistake", ["shifted_index", "dropped_last", "unsigned_read", "transposed_lanes"])
def test_planted_mistakes_are_caught(orc, mistake):
    m_vec, K, N, V = 8, 32, 16, 4
    csr = small_mask(m_vec, K, V, 0.4, seed=29, empty_rows=False)
    a = synthetic.random_lanes(csr.nnz_vec * V, 8, True, seed=30)
    b = synthetic.random_lanes(K * N, 8, True, seed=31)
    good = dense_spmm(m_vec, V, K, N, csr.row_offsets, csr.col_indices, a, b)
    ro, ci, aa, bb = csr.row_offsets.copy(), csr.col_indices.copy(), a.copy(), b.copy()
    if mistake == "shifted_index":
        ci = (ci + 1) % K
    elif mistake == "dropped_last":
        aa = aa.reshape(-1, V)
        aa[ro[1:] - 1] = 0
        aa = aa.ravel()
    elif mistake == "unsigned_read":
        aa = aa & 0xFF
    elif mistake == "transposed_lanes":
        aa = aa.reshape(-1, V)[:, ::-1].ravel()
    bad = orc.spmm_plain(m_vec, V, N, ro, ci, aa, bb)
    assert not np.array_equal(bad, good)
