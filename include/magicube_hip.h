/*
 * magicube_hip.h — C ABI of the MI355X (gfx950) quantized vector-sparse
 * SpMM / SDDMM library.  Plain pointers and sizes only; every device pointer
 * is caller-allocated HBM; the library allocates nothing in these calls and
 * keeps no state between them (except the tile-config table, see
 * mc_spmm_set_tile_config).  All launches are asynchronous on `stream`
 * (0 = the null stream) and never synchronize, so they can be captured in a
 * hipGraph.
 *
 * Formats are those of the Magicube reference (paths relative to it):
 *   row_offsets     2*m_vec int32 (start, end) pairs; start is aligned to the
 *                   pair's row alignment (16 slots for 8-bit B, 32 for 4-bit B),
 *                   end = start + real nonzero vectors
 *                   (SpMM/SpMM/spmm_benchmark.cpp:177-201)
 *   column_indices  int32 per slot, -1 in padding slots; for 4-bit B shuffled
 *                   within groups of 8 as [0,2,4,6,1,3,5,7]
 *                   (SpMM/SpMM/spmm_benchmark.cpp:204-222)
 *   values          prepared A values (SpMM/SpMM/spmm_benchmark.cpp:238-377);
 *                   padding slots must hold zeros, as the reference writes them
 *   rhs_matrix      K x N row-major, packed 4 x int8 / 8 x int4 per int32,
 *                   lane 0 in the low bits
 *   output_matrix   M x N int32 row-major (M = m_vec * vec_length), every
 *                   element written
 * Integer results are the exact sums modulo 2^32.
 */
#ifndef MAGICUBE_HIP_H
#define MAGICUBE_HIP_H

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum mc_status {
    MC_OK = 0,
    MC_EINVAL_V = 1,       /* vec_length not in {2, 4, 8} */
    MC_EINVAL_SHAPE = 2,   /* sizes out of range or not aligned as required */
    MC_EINVAL_ARG = 3,     /* null pointer, bad enum or bad config */
    MC_EHIP = 4,           /* a HIP runtime call failed: mc_last_hip_error() */
    MC_ENOMEM = 5,         /* caller-provided workspace too small */
    MC_EUNSUPPORTED = 6    /* combination not built */
} mc_status;

/* Lane interpretation of packed operands.  MC_SIGNED: two's complement
 * (int8 in [-128,127], int4 in [-8,7], int16 in [-32768,32767]).
 * MC_UNSIGNED: [0, 2^bits), what the reference's u8/u4 tensor-core MMAs
 * compute (SpMM/SpMM/src/spmm_utils/compute_utils.h:1431,1565). */
typedef enum mc_signedness { MC_SIGNED = 0, MC_UNSIGNED = 1 } mc_signedness;

/* Library metadata / errors. */
int mc_version(void);
const char *mc_status_string(int status);
int mc_last_hip_error(void);  /* hipError_t of the last MC_EHIP in this thread */

/* ---- SpMM ----------------------------------------------------------------
 * Replace spmm::wmmaSpmm_8b / _16b8b / _8b4b / _4b
 * (SpMM/SpMM/include/wmma_spmm.cuh:15-70, launchers
 * SpMM/SpMM/src/wmma_spmm.cu:2091-2383): same arguments in the same order,
 * plus the lane interpretation and a stream.  n must be a multiple of 4
 * (8-bit B) or 8 (4-bit B); k*n*bits(B)/8 < 2^30.  row_indices may be NULL
 * (identity order). */
mc_status mc_spmm_l8r8(int m_vec, int vec_length, int n, int k, const int *row_indices,
                       const int *row_offsets, const int *column_indices, const int *values,
                       const int *rhs_matrix, int *output_matrix, mc_signedness sign,
                       hipStream_t stream);
mc_status mc_spmm_l16r8(int m_vec, int vec_length, int n, int k, const int *row_indices,
                        const int *row_offsets, const int *column_indices, const int *values,
                        const int *rhs_matrix, int *output_matrix, mc_signedness sign,
                        hipStream_t stream);
mc_status mc_spmm_l8r4(int m_vec, int vec_length, int n, int k, const int *row_indices,
                       const int *row_offsets, const int *column_indices, const int *values,
                       const int *rhs_matrix, int *output_matrix, mc_signedness sign,
                       hipStream_t stream);
mc_status mc_spmm_l4r4(int m_vec, int vec_length, int n, int k, const int *row_indices,
                       const int *row_offsets, const int *column_indices, const int *values,
                       const int *rhs_matrix, int *output_matrix, mc_signedness sign,
                       hipStream_t stream);

/* Tile configurations of the SpMM kernel (an index into a fixed table):
 * lane_dwords B dwords gathered per lane per row, tiles_per_block column
 * tiles per workgroup, split_k waves sharing one (row, tile) over the
 * nonzero vectors, depth k-blocks of gathers in flight per wave.  No
 * reference counterpart (the reference fixes its tile
 * per template, SpMM/SpMM/src/wmma_spmm.cu:2137-2153); used by the
 * autotuner. */
typedef struct mc_tile_config {
    int lane_dwords;
    int tiles_per_block;
    int split_k;
    int depth;
} mc_tile_config;
int mc_spmm_num_tile_configs(void);
mc_status mc_spmm_get_tile_config(int index, mc_tile_config *out);
/* Generic SpMM with an explicit config index (-1: the table default for the
 * shape).  lhs_bits in {4, 8, 16}, rhs_bits in {4, 8}. */
mc_status mc_spmm_run(int lhs_bits, int rhs_bits, mc_signedness sign, int config, int m_vec,
                      int vec_length, int n, int k, const int *row_indices,
                      const int *row_offsets, const int *column_indices, const int *values,
                      const int *rhs_matrix, int *output_matrix, hipStream_t stream);
/* Installs the config the named entry points use for (lhs_bits, rhs_bits,
 * vec_length, n) — the autotuner's result; -1 restores the default. */
mc_status mc_spmm_set_tile_config(int lhs_bits, int rhs_bits, int vec_length, int n, int config);

/* ---- SDDMM ----------------------------------------------------------------
 * Replace sddmm::wmmaSddmm_8b / _4b (SDDMM/SDDMM/include/wmma_sddmm.cuh:7-32,
 * launchers SDDMM/SDDMM/src/wmma_sddmm.cu:376-465): same arguments in the
 * same order (note k before n), plus the lane interpretation and a stream.
 *   out[j][v] = sum_k lhs[(i*V+v)][k] * rhs[col_j][k]
 * lhs is (m_vec*vec_length) x k, rhs n x k, both row-major packed; k is 64,
 * 128 or 256; row starts are multiples of 8.  Every slot of
 * [start, start + roundup8(end - start)) is written, padding slots with 0, in
 * the reference layout [(j/8)][v][j%8] (SDDMM/SDDMM/sddmm_benchmark.cpp:55). */
mc_status mc_sddmm_l8r8(int m_vec, int k, int n, const int *row_indices, const int *row_offsets,
                        const int *col_indices, const int *lhs_matrix, const int *rhs_matrix,
                        int *output_values, int vec_length, mc_signedness sign, hipStream_t stream);
mc_status mc_sddmm_l4r4(int m_vec, int k, int n, const int *row_indices, const int *row_offsets,
                        const int *col_indices, const int *lhs_matrix, const int *rhs_matrix,
                        int *output_values, int vec_length, mc_signedness sign, hipStream_t stream);
/* Output layouts of mc_sddmm_run. */
#define MC_SDDMM_OUT_REFERENCE 0   /* [(j/8)][v][j%8] */
#define MC_SDDMM_OUT_VECTOR 1      /* [j][v] (vector-major, the attention layout) */
/* Generic SDDMM: bits in {4, 8} for both operands, explicit output layout. */
mc_status mc_sddmm_run(int bits, mc_signedness sign, int layout, int m_vec, int k, int n,
                       const int *row_indices, const int *row_offsets, const int *col_indices,
                       const int *lhs_matrix, const int *rhs_matrix, int *output_values,
                       int vec_length, hipStream_t stream);

/* ---- Host-side format preparation (no device work) ----------------------
 * Builds the reference layouts from a plain vector-CSR (row_offsets has
 * m_vec+1 entries, col_indices nnz_vec entries, values nnz_vec packed
 * vectors of vec_length lanes of lhs_bits each, lane 0 in the low bits).
 * Mirrors the preparation in BmFN (SpMM/SpMM/spmm_benchmark.cpp:177-377). */

/* Row alignment of a pair: 32 when either operand is 4-bit, else 16
 * (SpMM/SpMM/spmm_benchmark.cpp:140-143). */
int mc_format_row_alignment(int lhs_bits, int rhs_bits);
/* Writes 2*m_vec (start, end) pairs; returns the padded slot count
 * (aligned_num_item) or -1 on bad input. */
int mc_format_align_rows(int m_vec, const int *row_offsets, int alignment, int *aligned_pairs);
/* Padded column indices (-1 in padding), shuffled within 8-groups when
 * `shuffle` (4-bit B). */
mc_status mc_format_align_cols(int m_vec, const int *row_offsets, const int *col_indices,
                               const int *aligned_pairs, int aligned_num, int shuffle,
                               int *aligned_cols);
/* Prepared A values: aligned_num*vec_length*lhs_bits/8 bytes into `out`. */
mc_status mc_format_prepare_values(int lhs_bits, int rhs_bits, int vec_length, int m_vec,
                                   const int *row_offsets, const int *aligned_pairs,
                                   int aligned_num, const void *values, void *out);
/* Row order longest first, ties by row id (SpMM/SpMM/include/bm_test_utils.h:222-237). */
mc_status mc_format_sorted_rows(int m_vec, const int *row_offsets, int *row_indices);

#ifdef __cplusplus
}
#endif

#endif /* MAGICUBE_HIP_H */
