/*
 * magicube_vendor.h — vendor-library comparators for the SpMM / SDDMM
 * benchmarks (libmagicube_vendor.so).
 *
 * The reference measures its kernels against cuSPARSE Blocked-ELL SpMM with
 * V x V blocks in int8 and fp16 (baselines/spmm_benchmark.cpp:268-602), cuSPARSE
 * CSR SpMM (baselines/spmm_benchmark.cpp:162-219), cuSPARSE SDDMM
 * (baselines/sddmm_benchmark.cpp:33-88), cuBLAS int8 GEMM
 * (baselines/spmm_benchmark.cpp:657-668, baselines/sddmm_benchmark.cpp:362-384)
 * and cuBLAS fp16 GEMM (baselines/src/cublas_gemm.cu:19-28).  On MI355X the same
 * roles are played by rocSPARSE's generic SpMM and SDDMM and by hipBLASLt.
 * rocSPARSE's Blocked-ELL SpMM has no int8 or fp16 types, so V x V blocks go
 * through its BSR SpMM, which has both.  These entry points are benchmark infrastructure, not part of the
 * drop-in boundary: they take the vendor formats (scalar CSR, BSR, dense
 * row-major), which the caller builds from the vector-sparse operands.
 *
 * A plan owns the library handle, descriptors and workspace (allocated at
 * creation); mcv_run() only launches, on the plan's stream.  The operand
 * pointers are bound at creation and must stay valid until mcv_destroy().
 */
#ifndef MAGICUBE_VENDOR_H
#define MAGICUBE_VENDOR_H

#include <stdint.h>
#include <hip/hip_runtime_api.h>

#include "magicube_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum mcv_dtype {
    MCV_I8 = 0,   /* int8 operands, int32 accumulation and output (exact) */
    MCV_F16 = 1   /* fp16 operands, fp32 accumulation; output noted per call */
} mcv_dtype;

/* rocSPARSE CSR SpMM algorithms (rocsparse_spmm_alg). */
typedef enum mcv_csr_alg {
    MCV_CSR_DEFAULT = 0,
    MCV_CSR_ROW_SPLIT = 1,
    MCV_CSR_NNZ_SPLIT = 2,
    MCV_CSR_MERGE_PATH = 3
} mcv_csr_alg;

typedef struct mcv_plan mcv_plan;

/* C[m x n] = A[m x k] * B[k x n] with A in scalar CSR (m + 1 row offsets,
 * nnz column indices and values, zero-based), B and C dense row-major.
 * MCV_I8: C int32.  MCV_F16: C fp32.  Replaces the cuSPARSE CSR call of
 * baselines/spmm_benchmark.cpp:162-219 (CUSPARSE_SPMM_CSR_ALG2). */
mc_status mcv_csr_spmm_create(mcv_dtype dtype, mcv_csr_alg alg, int64_t m, int64_t k, int64_t n,
                              int64_t nnz, const int *row_ptr, const int *col_ind, const void *values,
                              const void *b, void *c, hipStream_t stream, mcv_plan **plan);

/* The same product with A in BSR: mb block rows, kb block columns, nnzb
 * blocks of block_dim x block_dim values, each block row-major.  Types as
 * for mcv_csr_spmm_create.  Stands for the cuSPARSE Blocked-ELL SpMM
 * (CUSPARSE_SPMM_BLOCKED_ELL_ALG1) of baselines/spmm_benchmark.cpp:268-602. */
mc_status mcv_bsr_spmm_create(mcv_dtype dtype, int64_t mb, int64_t kb, int64_t n, int64_t nnzb,
                              int block_dim, const int *row_ptr, const int *col_ind, const void *values,
                              const void *b, void *c, hipStream_t stream, mcv_plan **plan);

/* SDDMM: out_values[j] = sum_t lhs[r, t] * rhs[c, t] for each stored entry j
 * (row r, column c) of an m x n scalar CSR pattern; lhs m x k and rhs n x k
 * row-major fp16, out_values fp32 (rocSPARSE SDDMM has no integer types).
 * Stands where baselines/sddmm_benchmark.cpp:33-88 calls cuSPARSE SDDMM. */
mc_status mcv_csr_sddmm_create(int64_t m, int64_t n, int64_t k, int64_t nnz, const int *row_ptr,
                               const int *col_ind, const void *lhs, const void *rhs, void *out_values,
                               hipStream_t stream, mcv_plan **plan);

/* hipBLASLt dense GEMM: C[m x n] = A[m x k] * B, A row-major.  B is k x n
 * row-major (b_transposed = 0) or stored as its n x k transpose, row-major
 * (b_transposed = 1).  MCV_I8: C int32; MCV_F16: C fp16.  Creation times up to
 * eight of the library's heuristic candidates and keeps the fastest.  Replaces
 * cublasGemmEx (baselines/spmm_benchmark.cpp:657-668) and the fp16 GEMM of
 * baselines/src/cublas_gemm.cu:19-28. */
mc_status mcv_gemm_create(mcv_dtype dtype, int64_t m, int64_t k, int64_t n, const void *a, const void *b,
                          int b_transposed, void *c, hipStream_t stream, mcv_plan **plan);

/* Launches the planned operation on the plan's stream (asynchronous). */
mc_status mcv_run(mcv_plan *plan);

/* Workspace bytes the plan holds (0 if none). */
size_t mcv_workspace_bytes(const mcv_plan *plan);

/* Releases the plan (synchronizes its stream first). */
mc_status mcv_destroy(mcv_plan *plan);

/* Text of the last failure in this thread (library and status code). */
const char *mcv_last_error(void);

#ifdef __cplusplus
}
#endif

#endif  /* MAGICUBE_VENDOR_H */
