/*
 * magicube_hip.h — C ABI of the MI355X (gfx950) quantized vector-sparse
 * SpMM / SDDMM library.  Plain pointers and sizes only; every device pointer
 * is caller-allocated HBM; the library allocates nothing in these calls and
 * keeps no state between them.  All launches are asynchronous on `stream`
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
 * Replace spmm::wmmaSpmm_8b / _16b8b / _12b8b / _16b / _8b4b / _12b4b /
 * _16b4b / _4b (SpMM/SpMM/include/wmma_spmm.cuh:15-78, launchers
 * SpMM/SpMM/src/wmma_spmm.cu:2091-2459): same arguments in the same order,
 * plus the lane interpretation and a stream.  n must be a multiple of 2
 * (16-bit B), 4 (8-bit B) or 8 (4-bit B); k*n*bits(B)/8 < 2^30.
 * row_indices may be NULL (identity order).  12-bit A lanes sit in 16-bit
 * containers (the reference's preA = 16, preA_cut = 12); bits 12..15 are
 * ignored. */
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
/* spmm::wmmaSpmm_12b8b, SpMM/SpMM/include/wmma_spmm.cuh:55-61, wmma_spmm.cu:2385-2412 */
mc_status mc_spmm_l12r8(int m_vec, int vec_length, int n, int k, const int *row_indices,
                        const int *row_offsets, const int *column_indices, const int *values,
                        const int *rhs_matrix, int *output_matrix, mc_signedness sign,
                        hipStream_t stream);
/* spmm::wmmaSpmm_16b (16-bit A and B), wmma_spmm.cuh:71-77, wmma_spmm.cu:2434-2459 */
mc_status mc_spmm_l16r16(int m_vec, int vec_length, int n, int k, const int *row_indices,
                         const int *row_offsets, const int *column_indices, const int *values,
                         const int *rhs_matrix, int *output_matrix, mc_signedness sign,
                         hipStream_t stream);
/* spmm::wmmaSpmm_12b4b, wmma_spmm.cuh:31-37, wmma_spmm.cu:2208-2235 */
mc_status mc_spmm_l12r4(int m_vec, int vec_length, int n, int k, const int *row_indices,
                        const int *row_offsets, const int *column_indices, const int *values,
                        const int *rhs_matrix, int *output_matrix, mc_signedness sign,
                        hipStream_t stream);
/* spmm::wmmaSpmm_16b4b, wmma_spmm.cuh:39-45, wmma_spmm.cu:2260-2287 */
mc_status mc_spmm_l16r4(int m_vec, int vec_length, int n, int k, const int *row_indices,
                        const int *row_offsets, const int *column_indices, const int *values,
                        const int *rhs_matrix, int *output_matrix, mc_signedness sign,
                        hipStream_t stream);

/* Tile configurations of the SpMM kernels (an index into a fixed table).
 * kind 0, gather kernel: lane_dwords B dwords gathered per lane per row,
 *   tiles_per_block column tiles per workgroup, split_k waves sharing one
 *   (row, tile) over the nonzero vectors, depth k-blocks of gathers in
 *   flight per wave.
 * kind 1, B-stationary kernel (8-bit B; the K rows of a column slice of B
 *   are held in LDS): lane_dwords B dwords per row of the slice (n must be a
 *   multiple of 4*lane_dwords and (k+1)*4*lane_dwords bytes must fit in LDS),
 *   tiles_per_block waves per workgroup, split_k vector rows per workgroup,
 *   depth k-blocks of indices and A operands in flight per wave.
 * waves_per_simd > 1: the kernel's registers are allocated so that this many
 *   waves fit on a SIMD (0: the compiler's choice).
 * No reference counterpart (the reference fixes its tile per template,
 * SpMM/SpMM/src/wmma_spmm.cu:2137-2153); used by the autotuner. */
typedef struct mc_tile_config {
    int lane_dwords;
    int tiles_per_block;
    int split_k;
    int depth;
    int kind;
    int waves_per_simd;
} mc_tile_config;
int mc_spmm_num_tile_configs(void);
mc_status mc_spmm_get_tile_config(int index, mc_tile_config *out);
/* Generic SpMM with an explicit config index (-1: the default for the shape:
 * one dword of B per lane, up to four waves per row as n allows, and the
 * row's k-blocks split over two waves when the launch would hold fewer than
 * 2048 waves).  (lhs_bits, rhs_bits) in {(8,8), (12,8), (16,8), (16,16),
 * (4,4), (8,4), (12,4), (16,4)}. */
mc_status mc_spmm_run(int lhs_bits, int rhs_bits, mc_signedness sign, int config, int m_vec,
                      int vec_length, int n, int k, const int *row_indices,
                      const int *row_offsets, const int *column_indices, const int *values,
                      const int *rhs_matrix, int *output_matrix, hipStream_t stream);
/* Batched SpMM over `batch` matrices sharing one mask (row_indices,
 * row_offsets, column_indices): matrix b's prepared values start
 * b * values_stride_bytes after `values`, its B b * rhs_stride_bytes after
 * `rhs_matrix` and its C b * out_stride int32 after `output_matrix`.  One
 * launch, matrices in gridDim.y (batch <= 65535; byte strides multiples of
 * 4, MC_EINVAL_SHAPE otherwise); gather tile configs only
 * (B-stationary configs return MC_EUNSUPPORTED).  Replaces the batched
 * launchers behind the reference's bspmm_8b / _16b8b / _8b4b / _4b
 * (end2end_eval/sparse_transformer_magicube/src/cuda/deq_spmm.cpp:5-141,
 * deq_spmm_kernel.cu:1287-1361), which launch the batch as one grid. */
mc_status mc_spmm_batched(int lhs_bits, int rhs_bits, mc_signedness sign, int config, int batch, int m_vec,
                          int vec_length, int n, int k, const int *row_indices, const int *row_offsets,
                          const int *column_indices, const int *values, long long values_stride_bytes,
                          const int *rhs_matrix, long long rhs_stride_bytes, int *output_matrix,
                          long long out_stride, hipStream_t stream);

/* ---- SDDMM ----------------------------------------------------------------
 * Replace sddmm::wmmaSddmm_8b / _4b / _16b (SDDMM/SDDMM/include/wmma_sddmm.cuh:7-32,
 * launchers SDDMM/SDDMM/src/wmma_sddmm.cu:376-514): same arguments in the
 * same order (note k before n), plus the lane interpretation and a stream.
 *   out[j][v] = sum_k lhs[(i*V+v)][k] * rhs[col_j][k]
 * lhs is (m_vec*vec_length) x k, rhs n x k, both row-major packed; k is a
 * multiple of 64 (as the reference's Tile_K loop); row starts are multiples
 * of 8.  Every slot of
 * [start, start + roundup8(end - start)) is written, padding slots with 0, in
 * the reference layout [(j/8)][v][j%8] (SDDMM/SDDMM/sddmm_benchmark.cpp:55). */
mc_status mc_sddmm_l8r8(int m_vec, int k, int n, const int *row_indices, const int *row_offsets,
                        const int *col_indices, const int *lhs_matrix, const int *rhs_matrix,
                        int *output_values, int vec_length, mc_signedness sign, hipStream_t stream);
mc_status mc_sddmm_l4r4(int m_vec, int k, int n, const int *row_indices, const int *row_offsets,
                        const int *col_indices, const int *lhs_matrix, const int *rhs_matrix,
                        int *output_values, int vec_length, mc_signedness sign, hipStream_t stream);
/* sddmm::wmmaSddmm_16b (16-bit lhs and rhs), wmma_sddmm.cuh:25-32 */
mc_status mc_sddmm_l16r16(int m_vec, int k, int n, const int *row_indices, const int *row_offsets,
                          const int *col_indices, const int *lhs_matrix, const int *rhs_matrix,
                          int *output_values, int vec_length, mc_signedness sign, hipStream_t stream);
/* Output layouts of mc_sddmm_run. */
#define MC_SDDMM_OUT_REFERENCE 0   /* [(j/8)][v][j%8] */
#define MC_SDDMM_OUT_VECTOR 1      /* [j][v] (vector-major, the attention layout) */
/* Generic SDDMM: bits in {4, 8, 16} for both operands, explicit output layout. */
mc_status mc_sddmm_run(int bits, mc_signedness sign, int layout, int m_vec, int k, int n,
                       const int *row_indices, const int *row_offsets, const int *col_indices,
                       const int *lhs_matrix, const int *rhs_matrix, int *output_values,
                       int vec_length, hipStream_t stream);
/* Batched SDDMM over `batch` matrices sharing one mask: matrix b's lhs and
 * rhs start b * lhs_stride_bytes / b * rhs_stride_bytes after the first,
 * its output b * out_stride int32 after.  One launch (batch <= 65535);
 * byte strides are multiples of 4 (MC_EINVAL_SHAPE otherwise).
 * Replaces the batched launcher behind the reference's bsddmm_8b / _4b
 * (end2end_eval/sparse_transformer_magicube/src/cuda/deq_sddmm.cpp:56-93,
 * deq_sddmm_kernel.cu:565-636). */
mc_status mc_sddmm_batched(int bits, mc_signedness sign, int layout, int batch, int m_vec, int k, int n,
                           const int *row_indices, const int *row_offsets, const int *col_indices,
                           const int *lhs_matrix, long long lhs_stride_bytes, const int *rhs_matrix,
                           long long rhs_stride_bytes, int *output_values, long long out_stride, int vec_length,
                           hipStream_t stream);

/* ---- Quantized sparse attention ------------------------------------------
 * Layer semantics of spMultiheadAttention
 * (end2end_eval/sparse_transformer_magicube/spattention.py:182-216) per
 * (sample, head) matrix of sequence length L and head dim 64, mask rows
 * aligned to 16 slots:
 *   Qq, Kq, Vq = trunc(clamp(x * scale))          (quantization_kernel.cu:46-74)
 *                or rint(clamp(x * scale)) with MC_ROUND_NEAREST; the clamp
 *                range is that of the output width, or [-8, 7] for int8
 *                output with MC_QUANT_RANGE4 (4-bit values stored one per
 *                byte; the fused launch also takes 4-bit V packed, v_bits 4)
 *   S = Qq Kq^T at the mask (int32, exact)
 *   s = float(S) * alpha     alpha = 1 / (scale_q * scale_k * sqrt(64))
 *   P, O by the requantization mode:
 *     MC_P_REFERENCE  P = trunc(min(exp(s - max) / sum * p_scale, 127)),
 *                     O = float(P Vq) * o_scale           (q_softmax_kernel.cu:132-147)
 *     MC_P_ROW_SCALED P = rint(min(exp(s - max) * p_scale, 127)),
 *                     O = (float(P Vq) * o_scale) * (1 / sum)
 *     MC_P_REFERENCE_4BIT  as MC_P_REFERENCE with P clamped at 7: the
 *                     reference's 4-bit P (q_softmax_kernel.cu:84-118, lhs_pre 4),
 *                     written one value per byte where P is an output
 * The softmax statistics are accumulated over 64-slot chunks in a fixed
 * order, shared by every kernel below, so the one-op-per-launch path and the
 * fused batched path produce identical int32 S, int8 P and int32 P Vq. */
#define MC_P_REFERENCE 0
#define MC_P_ROW_SCALED 1
#define MC_P_REFERENCE_4BIT 2
#define MC_ROUND_TRUNC 0
#define MC_ROUND_NEAREST 1
/* flag, or-ed into the rounding argument of mc_attn_quantize_batched */
#define MC_QUANT_RANGE4 2

/* One quantization job: count fp16 values -> int8 bytes or packed int4.
 * row_len == 0: x holds the count values contiguously.  Otherwise x holds
 * rows of row_len values (a multiple of 8) row_stride values apart, e.g. one
 * head's [L][64] slice of a [L][batch][3 * heads * 64] projection output;
 * out is always contiguous. */
typedef struct mc_quant_job {
    const void *x;
    long long count;
    float scale;
    void *out;
    long long row_len;
    long long row_stride;
} mc_quant_job;
/* Replaces sptrans bquantization (src/cuda/quantization.cpp:3-20): jobs is a
 * device array of njobs jobs; max_count bounds every job's count. */
mc_status mc_attn_quantize_batched(const mc_quant_job *jobs, int njobs, int bits, int rounding,
                                   long long max_count, hipStream_t stream);
/* max |x| of an fp16 tensor, as float bits combined into *out_bits (device,
 * zero-initialised by the caller) with atomicMax.  No reference counterpart
 * (the reference uses a fixed scale, spattention.py:51). */
mc_status mc_attn_absmax(const void *x_fp16, long long count, unsigned *out_bits, hipStream_t stream);
/* Replaces q_bcsr_softmax (src/cuda/q_softmax.cpp:26-52) for one matrix:
 * s_vals is the int32 SDDMM output in vector-major layout [slot][V]; P is
 * written in the SpMM A layout (16-slot groups of [V][16] bytes); row_inv_l
 * (optional, m_vec*V floats) receives 1/sum per row. */
mc_status mc_attn_softmax_requant(int m_vec, int vec_length, const int *row_indices,
                                  const int *row_offsets, const int *s_vals, float alpha,
                                  float p_scale, int p_mode, void *p_out, float *row_inv_l,
                                  hipStream_t stream);
/* out[r][c] = (float(in[r][c]) * scale) * row_scale[r] (row_scale may be NULL):
 * the dequantizing epilogue of the reference's bspmm (deq_spmm_kernel.cu:645-670). */
mc_status mc_attn_dequant(const int *in, int rows, int cols, float scale, const float *row_scale,
                          float *out, hipStream_t stream);
/* out_f16[i] = half(float(in[i]) / divisor), round to nearest even: the
 * reference's fp16 epilogues of bsddmm (sddmm_utils/output_tile.h:37-43) and
 * bspmm (deq_spmm_kernel.cu:645-670). */
mc_status mc_attn_dequant_f16(const int *in, long long count, float divisor, void *out_f16,
                              hipStream_t stream);
/* The reference's q_bcsr_softmax (src/cuda/q_softmax.cpp:26-52,
 * q_softmax_kernel.cu:180-271,394-479) on fp16 logits, for `batch` matrices
 * sharing one mask: values_f16 holds matrix b's [slot][V] logits at
 * b * values_stride halves; vector row m = row_indices[x] (x < rows; NULL:
 * x) gets
 *   s = float(logit) * sqrt_dk,  max_v from -1e10 (v = 0) or 0 (v >= 1),
 *   P = (int) clamp(exp(s - max_v) / sum_v * scale, lo, hi)
 * at bits 16 / 8 / 4 in the reference's layouts, matrix b at attn +
 * b * attn_stride_bytes + start * V * bits / 8:
 *   8-bit  [(i/16)][v][16] bytes; 16-bit [(i/16)][v][low 16][high 16];
 *   4-bit  [(i/32)][v][16] bytes of nibble pairs (slot 2j low, 2j+1 high).
 * Every slot of [start, start + roundup(nonzeros, 32 for 4-bit, else 16))
 * is written (padding with 0).  vec_length 8 needs 16-byte aligned logits
 * (values_f16 and values_stride a multiple of 8). */
mc_status mc_attn_softmax_f16(int batch, int rows, int vec_length, const int *row_indices,
                              const int *row_offsets, const void *values_f16, long long values_stride,
                              float sqrt_dk, float scale, int bits, void *attn, long long attn_stride_bytes,
                              hipStream_t stream);

/* One attention matrix of the fused batched launch (device-resident
 * descriptor).  q, k: [L][64] int8 (or int4 packed [L][32] bytes); v: [L][64]
 * int8, or int4 packed [L][32] bytes with v_bits 4; o: [L][64] float.  o_int (int32 P Vq), p_out (P in the SpMM A layout)
 * and s_out (int32 S, [slot][V]) are optional debug outputs.  L <= 2^24 (the
 * kernel addresses rows with 32-bit byte offsets and marks masked slots with
 * offsets >= 2^31). */
typedef struct mc_attn_matrix {
    int L;
    int m_vec;
    const int *row_offsets;
    const int *col_indices;
    const void *q;
    const void *k;
    const void *v;
    float alpha;
    float o_scale;
    float *o;
    int *o_int;
    unsigned char *p_out;
    int *s_out;
    void *o16;         /* non-null: fp16 output (round to nearest) at o16 +   */
    int o16_stride;    /* row * o16_stride + column (a multiple of 4, o16     */
    int o16_pad;       /* 8-byte aligned) instead of the fp32 o               */
    const void *q16;   /* non-null: Q as fp16 rows at q16 + row * q16_stride  */
    int q16_stride;    /* (q16 16-byte aligned, q16_stride a multiple of 8,   */
    float q_scale;     /* (L - 1) * q16_stride + 64 < 2^30), quantized in the */
    int q_rounding;    /* kernel as mc_attn_quantize_batched quantizes with   */
    int q_pad;         /* q_scale and q_rounding; q is then not read          */
} mc_attn_matrix;
/* Fused quantized sparse attention over many matrices in one persistent
 * launch.  items is a device array of n_items (matrix index, vector row)
 * int pairs (pairs with a negative matrix index are skipped), processed by
 * grid_waves waves (0: one wave per item) in list order with a stride of
 * grid_waves; any order gives the same results.  Item i runs in workgroup
 * i / 4, and consecutive workgroups go to the 8 XCDs in turn, so an order
 * that keeps each matrix on one XCD (see the Python package) keeps its K/V
 * rows in that XCD's L2.
 * (qk_bits, v_bits) is (8, 8), (4, 8) or (4, 4): 4-bit V packed as the
 * reference's rhs_pre 4 layer holds it (its bspmm_8b4b / bspmm_4b operand,
 * deq_spmm.cpp:50-118); MC_EUNSUPPORTED otherwise.
 * debug != 0 also writes the descriptors' non-null o_int / p_out / s_out
 * (a slower kernel); with debug == 0 they are ignored.  No reference
 * counterpart: the reference launches each op separately with one mask
 * shared by the batch (spattention.py:189-210). */
mc_status mc_attn_forward(const mc_attn_matrix *matrices, const int *items, int n_items, int qk_bits,
                          int v_bits, int vec_length, float p_scale, int p_mode, int grid_waves, int debug,
                          hipStream_t stream);

/* LayerNorm over rows of cols fp16 values (cols 256, 512 or 1024), fused
 * with the residual add before it when residual is non-null:
 *   s = half(x + residual),  out = half((s - mean) * rsqrt(var + eps) * weight + bias)
 * mean and the biased variance of the row in fp32; weight / bias (fp16, cols
 * values) may be null (1 and 0).  x, residual and out 8-byte aligned; out may
 * alias x or residual.  No counterpart in the reference's library: its
 * encoder uses torch's LayerNorm after a separate add
 * (end2end_eval/sparse_transformer_magicube/end_to_end.py:77-89). */
mc_status mc_layernorm_f16(long long rows, int cols, const void *x, const void *residual, const void *weight,
                           const void *bias, float eps, void *out, hipStream_t stream);

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
/* Prepared A values: aligned_num*vec_length*c/8 bytes into `out`, where c is
 * the lane container width (16 for 12-bit lanes, else lhs_bits); `values`
 * holds nnz vectors of vec_length lanes at that width. */
mc_status mc_format_prepare_values(int lhs_bits, int rhs_bits, int vec_length, int m_vec,
                                   const int *row_offsets, const int *aligned_pairs,
                                   int aligned_num, const void *values, void *out);
/* Row order longest first, ties by row id (SpMM/SpMM/include/bm_test_utils.h:222-237). */
mc_status mc_format_sorted_rows(int m_vec, const int *row_offsets, int *row_indices);

#ifdef __cplusplus
}
#endif

#endif /* MAGICUBE_HIP_H */
