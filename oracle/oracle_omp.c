/*
 * Row-parallel variant of the SpMM host check (TEST INFRASTRUCTURE ONLY: the
 * cpu_baseline leg of bench.py times it beside the single-thread loop).
 * Same arithmetic as orc_spmm_ref_check (compute_ref_integers,
 * SpMM/SpMM/spmm_benchmark.cpp:69-120) for scaleA = 1 and preA_cut = preA:
 * an A lane is the unsigned preA-bit field of the TypeA container (the
 * literal check masks it to preA bits after the shift, :101), a B lane is
 * extracted in int arithmetic, so the top lane of each int32 sign-extends
 * (:106).  Vector rows write disjoint output rows, so OpenMP splits the outer
 * loop with no reduction on C.  tests/test_oracle.py pins it bit for bit
 * against orc_spmm_ref_check.
 */
#include "oracle.h"

#include <omp.h>
#include <string.h>

static int32_t omp_extract_int(int32_t tile, int pre, int shift) {
    int32_t m = (int32_t)((uint32_t)((1u << pre) - 1u) << shift);
    return (m & tile) >> shift;
}

static uint64_t omp_load_container(const void *A, int type_bytes, long j) {
    switch (type_bytes) {
    case 1: return ((const uint8_t *)A)[j];
    case 2: return ((const uint16_t *)A)[j];
    case 4: return ((const uint32_t *)A)[j];
    default: return ((const uint64_t *)A)[j];
    }
}

double orc_spmm_ref_check_omp(const void *A, int type_bytes, const int32_t *B, int32_t *C,
                              int M, int K, int N, int preA, int preB, int V,
                              const int *row_offsets, const int *col_indices, int m_vec,
                              int threads) {
    (void)K;
    if (threads > 0) omp_set_num_threads(threads);
    memset(C, 0, sizeof(int32_t) * (size_t)M * (size_t)N);
    const int b_tile = 32 / preB;
    const uint64_t maskA = (1ULL << preA) - 1ULL;
    double flops = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : flops)
    for (int i = 0; i < m_vec; i++) {
        for (int j = row_offsets[i]; j < row_offsets[i + 1]; j++) {
            int col_idx = col_indices[j];
            uint64_t tile = omp_load_container(A, type_bytes, j);
            for (int av = 0; av < V; av++) {
                int a_val = (int)((tile >> (av * preA)) & maskA);
                uint32_t *crow = (uint32_t *)C + (long)(i * V + av) * N;
                const int32_t *brow = B + (long)col_idx * (N / b_tile);
                for (int n = 0; n < N / b_tile; n++) {
                    int32_t Bt = brow[n];
                    for (int bv = 0; bv < b_tile; bv++)
                        crow[n * b_tile + bv] += (uint32_t)(a_val * omp_extract_int(Bt, preB, bv * preB));
                }
                flops += 2.0 * N;
            }
        }
    }
    return flops;
}
