/*
 * CPU oracle for the quantized vector-sparse SpMM / SDDMM path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (magicube_amd/, include/)
 * links, loads or calls this code; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, and only as the checker.
 *
 * Every function restates one piece of the reference (Shigangli/Magicube) in
 * plain C.  Citations are path:line relative to the reference checkout:
 *
 *   orc_make_dense_matrix      SpMM/SpMM/include/bm_test_utils.h:130-164
 *                              (SDDMM copy: SDDMM/SDDMM/include/bm_test_utils.h:128-171)
 *   orc_align_row_offsets      SpMM/SpMM/spmm_benchmark.cpp:177-201
 *                              (SDDMM: SDDMM/SDDMM/sddmm_benchmark.cpp:104-115)
 *   orc_align_col_indices      SpMM/SpMM/spmm_benchmark.cpp:204-216
 *   orc_shuffle_col_indices    SpMM/SpMM/spmm_benchmark.cpp:218-222
 *   orc_prepare_values         SpMM/SpMM/spmm_benchmark.cpp:238-377, 425-428
 *   orc_spmm_ref_check         SpMM/SpMM/spmm_benchmark.cpp:69-120 (compute_ref_integers)
 *   orc_sddmm_ref_check        SDDMM/SDDMM/sddmm_benchmark.cpp:24-60 (Host_sddmm_integers)
 *   orc_sorted_row_swizzle     SpMM/SpMM/include/bm_test_utils.h:222-237
 *
 * The "plain" functions (orc_spmm_plain, orc_sddmm_plain, orc_unpack_lanes)
 * state the same sums on unpacked integers, with the lane interpretation
 * (signed two's complement or unsigned) as an explicit argument.  All sums are
 * exact modulo 2^32 (int32 wrap-around), which is what the reference's int
 * accumulation produces on every input where it is defined.
 *
 * Pinning (see tests/test_oracle.py): the plain sums are checked against an
 * independent numpy dense matmul; the reference-literal checks are checked
 * against the plain sums on the reference's own generator data; the generator
 * restatement is checked against libstdc++'s <random> compiled with g++.
 */
#ifndef MC_ORACLE_H
#define MC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- reference data generator ------------------------------------------ */

/* Fills `count` elements of a TypeA array (type_bytes = 1 char, 2 short,
 * 4 int, 8 long long) exactly as MakeDenseMatrix<TypeA> does when handed a
 * fresh std::default_random_engine (std::minstd_rand0, seed 1) by value.
 * sddmm_variant selects the SDDMM copy of the header (long long differs). */
void orc_make_dense_matrix(int type_bytes, int64_t count, void *out, int sddmm_variant);

/* The raw canonical float stream (generate_canonical<float,24> over
 * minstd_rand0 seeded 1), for pinning against libstdc++. */
void orc_canonical_stream(int64_t count, float *out);

/* ---- vector-CSR preparation -------------------------------------------- */

/* Writes 2*m_vec (start, end) pairs; returns aligned_num_item. */
int orc_align_row_offsets(int m_vec, const int *row_offsets, int align, int *aligned_pairs);
void orc_align_col_indices(int m_vec, const int *row_offsets, const int *col_indices,
                           const int *aligned_pairs, int aligned_num, int *aligned_cols);
void orc_shuffle_col_indices(int aligned_num, const int *aligned_cols, int *shuffled);
void orc_sorted_row_swizzle(int rows, const int *row_offsets, int *row_indices);

/* Reference value preparation: `values` holds nnz_vec vectors of V lanes at
 * preA container bits (scaleA TypeA words per vector, TypeA of type_bytes).
 * Writes aligned_num*V*preA/8 bytes into `out`: the array BmFN copies to
 * d_values (the transposed buffer or the decomposed one, :425-428). */
void orc_prepare_values(const void *values, int type_bytes, int scaleA,
                        int m_vec, const int *row_offsets, const int *aligned_pairs,
                        int aligned_num, int V, int preA, int preA_cut, int preB,
                        void *out);

/* ---- reference host checks (literal) ------------------------------------ */

/* compute_ref_integers<TypeA>: A is the unpadded, untransposed `values`
 * array, B the packed K x N rhs_matrix.  Returns the "flops" count. */
double orc_spmm_ref_check(const void *A, int type_bytes, const int32_t *B, int32_t *C,
                          int M, int K, int N, int preA, int preA_cut, int preB, int V,
                          const int *row_offsets, const int *col_indices, int m_vec,
                          int scaleA);

/* Host_sddmm_integers: lhs M x K and rhs N x K packed int32; row_offsets are
 * the aligned (start,end) pairs; out uses the [(j/a)][v][j%a] layout and is
 * only written at real nonzeros. */
double orc_sddmm_ref_check(const int32_t *lhs, const int32_t *rhs, int32_t *out,
                           int M, int K, int N, int preA, int preB, int V,
                           const int *row_offsets, const int *col_indices, int m_vec,
                           int alignment);

/* ---- plain integer restatements ----------------------------------------- */

/* Extracts `count` lanes of `bits` bits from packed little-endian words
 * (lane 0 in the lowest bits).  is_signed: two's complement lanes. */
void orc_unpack_lanes(const void *packed, int bits, int64_t count, int is_signed, int32_t *out);
/* Inverse of orc_unpack_lanes (values are truncated to `bits`). */
void orc_pack_lanes(const int32_t *vals, int bits, int64_t count, void *packed);

/* C[(i*V+v)*N + n] = sum_j a[j*V+v] * b[col_j*N + n]  (mod 2^32) over the
 * plain CSR (row_offsets has m_vec+1 entries, a has nnz*V entries). */
void orc_spmm_plain(int m_vec, int V, int N, const int *row_offsets, const int *col_indices,
                    const int32_t *a, const int32_t *b, int32_t *C);

/* out[(j/a)*a*V + a*v + j%a] = sum_k lhs[(i*V+v)*K + k] * rhs[col_j*K + k] over
 * aligned pairs (real slots only; other slots untouched). */
void orc_sddmm_plain(int m_vec, int V, int K, const int *aligned_pairs, const int *aligned_cols,
                     const int32_t *lhs, const int32_t *rhs, int32_t *out, int alignment);

/* Row-parallel (OpenMP) copy of orc_spmm_ref_check for TypeA of 1, 2, 4 or
 * 8 bytes with scaleA = 1 and preA_cut = preA (oracle_omp.c); threads <= 0
 * keeps the default. */
double orc_spmm_ref_check_omp(const void *A, int type_bytes, const int32_t *B, int32_t *C,
                              int M, int K, int N, int preA, int preB, int V,
                              const int *row_offsets, const int *col_indices, int m_vec,
                              int threads);

#ifdef __cplusplus
}
#endif

#endif
