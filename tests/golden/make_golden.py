"""Writes tests/golden/refdata.npz: small SpMM and SDDMM cases on the
reference benchmarks' own data (MakeDenseMatrix restated,
SpMM/SpMM/include/bm_test_utils.h:130-164) with the expected outputs of the
literal host checks (compute_ref_integers, SpMM/SpMM/spmm_benchmark.cpp:69-120;
Host_sddmm_integers, SDDMM/SDDMM/sddmm_benchmark.cpp:24-60), computed by the
oracle.  Run from the repository root:  python tests/golden/make_golden.py

Arrays only (np.savez_compressed, no pickled objects).  Keys per case:
  spmm_L{a}R{b}_V{v}_{row_offsets,col_indices,values,rhs,out}
  sddmm_{b}b_V{v}_{row_offsets,col_indices,lhs,rhs,out}
row_offsets / col_indices are the plain pattern (m_vec + 1 offsets); values
and rhs hold the generator's words as bytes, lane 0 in the low bits (12-bit
lanes in the generator's 16-bit containers, whose bits 12..15 the checks drop).
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "refdata.npz")
# (preA, preB, V) -> (TypeA bytes, preA as generated, preA_cut, scaleA), SpMM/SpMM/spmm_benchmark.cpp:793-831
TYPES = {
    (8, 8, 8): (8, 8, 8, 1), (8, 8, 4): (4, 8, 8, 1), (8, 8, 2): (2, 8, 8, 1),
    (16, 8, 8): (8, 16, 16, 2), (16, 8, 4): (8, 16, 16, 1), (16, 8, 2): (4, 16, 16, 1),
    (8, 4, 8): (8, 8, 8, 1), (8, 4, 4): (4, 8, 8, 1), (8, 4, 2): (2, 8, 8, 1),
    (4, 4, 8): (4, 4, 4, 1), (4, 4, 4): (2, 4, 4, 1), (4, 4, 2): (1, 4, 4, 1),
    (12, 8, 8): (8, 16, 12, 2), (12, 8, 4): (8, 16, 12, 1), (12, 8, 2): (4, 16, 12, 1),
    (16, 16, 8): (2, 16, 16, 8), (16, 16, 4): (2, 16, 16, 4), (16, 16, 2): (2, 16, 16, 2),
    (12, 4, 8): (8, 16, 12, 2), (12, 4, 4): (8, 16, 12, 1), (12, 4, 2): (4, 16, 12, 1),
    (16, 4, 8): (8, 16, 16, 2), (16, 4, 4): (8, 16, 16, 1), (16, 4, 2): (4, 16, 16, 1),
}
M_VEC, K, N = 8, 64, 32


def pattern(m_vec, k, seed):
    """Deterministic vector-sparse pattern: rows of 0..k/4 sorted distinct columns, one empty row."""
    rng = np.random.default_rng(seed)
    counts = rng.integers(1, k // 4 + 1, size=m_vec)
    counts[m_vec // 2] = 0
    cols = [np.sort(rng.choice(k, size=c, replace=False)) for c in counts]
    ro = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    return ro, (np.concatenate(cols) if cols else np.zeros(0)).astype(np.int32)


def main(path: str = OUT):
    data = {}
    for (pa, pb, V), (tb, gen_a, cut, scale) in sorted(TYPES.items()):
        ro, ci = pattern(M_VEC, K, seed=100 + pa * 10 + pb + V)
        nnz = int(ro[-1])
        count = nnz * V * scale * gen_a // (tb * 8)
        vals = oracle.make_dense_matrix(tb, max(count, 1))
        rhs = oracle.make_dense_matrix(4, K * N * pb // 32)
        C, _ = oracle.spmm_ref_check(vals, tb, rhs, M_VEC * V, K, N, gen_a, cut, pb, V, ro, ci, M_VEC, scale)
        key = f"spmm_L{pa}R{pb}_V{V}"
        data[key + "_row_offsets"] = ro
        data[key + "_col_indices"] = ci
        data[key + "_values"] = vals.view(np.uint8)[: nnz * V * gen_a // 8].copy()
        data[key + "_rhs"] = rhs.view(np.uint8).copy()
        data[key + "_out"] = C
    for bits in (8, 4, 16):
        for V in (8, 4, 2):
            ro, ci = pattern(M_VEC, N, seed=200 + bits + V)
            pairs, an = oracle.align_row_offsets(ro, 8)
            cols = oracle.align_col_indices(ro, ci, pairs, an)
            lhs = oracle.make_dense_matrix(4, M_VEC * V * K * bits // 32, sddmm=True)
            rhs = oracle.make_dense_matrix(4, N * K * bits // 32, sddmm=True)
            out, _ = oracle.sddmm_ref_check(lhs, rhs, M_VEC * V, K, N, bits, bits, V, pairs, cols, M_VEC, 8,
                                            max(an, 8) * V)
            key = f"sddmm_{bits}b_V{V}"
            data[key + "_row_offsets"] = ro
            data[key + "_col_indices"] = ci
            data[key + "_lhs"] = lhs.view(np.uint8).copy()
            data[key + "_rhs"] = rhs.view(np.uint8).copy()
            data[key + "_out"] = out
    data["dims"] = np.array([M_VEC, K, N], dtype=np.int32)
    np.savez_compressed(path, **data)
    print(f"wrote {path}: {len(data)} arrays, {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main()
