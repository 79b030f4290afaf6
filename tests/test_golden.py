"""Golden fixtures (tests/golden/refdata.npz, written by
tests/golden/make_golden.py from the reference benchmarks' generator and
literal host checks): they regenerate bit-identically, and every expected
output equals an independent numpy computation on unsigned lanes."""
import importlib.util
import os

import numpy as np

from magicube_amd import formats

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load():
    return np.load(os.path.join(GOLDEN, "refdata.npz"), allow_pickle=False)


def cases(d, prefix):
    return sorted({k.rsplit("_", 1)[0] for k in d.files if k.startswith(prefix) and k.endswith("_out")})


def test_fixture_regenerates(tmp_path):
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(GOLDEN, "make_golden.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    out = tmp_path / "regen.npz"
    mod.main(str(out))
    a, b = load(), np.load(out, allow_pickle=False)
    assert sorted(a.files) == sorted(b.files)
    for k in a.files:
        assert np.array_equal(a[k], b[k]), k


def test_spmm_fixture_against_dense_numpy():
    d = load()
    m_vec, K, N = (int(x) for x in d["dims"])
    for key in cases(d, "spmm_"):
        pa, pb = (int(x) for x in key.split("_")[1][1:].split("R"))
        V = int(key.split("_V")[1])
        ro, ci = d[key + "_row_offsets"], d[key + "_col_indices"]
        a = formats.unpack_lanes(d[key + "_values"], pa, int(ro[-1]) * V, signed=False).astype(np.int64)
        b = formats.unpack_lanes(d[key + "_rhs"], pb, K * N, signed=False).astype(np.int64).reshape(K, N)
        A = np.zeros((m_vec * V, K), np.int64)
        for i in range(m_vec):
            for j in range(ro[i], ro[i + 1]):
                A[i * V:(i + 1) * V, ci[j]] += a[j * V:(j + 1) * V]
        exp = ((A @ b) & 0xFFFFFFFF).astype(np.uint32).view(np.int32)
        assert np.array_equal(d[key + "_out"], exp), key
        assert (d[key + "_out"] > 0).sum() > 0


def test_sddmm_fixture_against_dense_numpy():
    d = load()
    m_vec, K, N = (int(x) for x in d["dims"])
    for key in cases(d, "sddmm_"):
        bits = int(key.split("_")[1][:-1])
        V = int(key.split("_V")[1])
        ro, ci = d[key + "_row_offsets"], d[key + "_col_indices"]
        lhs = formats.unpack_lanes(d[key + "_lhs"], bits, m_vec * V * K, False).astype(np.int64).reshape(-1, K)
        rhs = formats.unpack_lanes(d[key + "_rhs"], bits, N * K, False).astype(np.int64).reshape(N, K)
        pairs, an = formats.align_rows(ro, 8)
        exp = np.zeros(max(an, 8) * V, np.int64)
        for i in range(m_vec):
            for t, j in enumerate(range(ro[i], ro[i + 1])):
                slot = pairs[2 * i] + t
                for v in range(V):
                    exp[(slot // 8) * 8 * V + 8 * v + slot % 8] = lhs[i * V + v] @ rhs[ci[j]]
        assert np.array_equal(d[key + "_out"], (exp & 0xFFFFFFFF).astype(np.uint32).view(np.int32)), key
