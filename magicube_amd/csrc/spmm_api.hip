// C ABI of the SpMM family: argument validation, the tile-config table and
// dispatch to the per-pair instantiation units.  The entry points keep no
// state: the tile is the default for the shape unless the call names one.
#include "spmm_dispatch.hpp"

namespace mc {
namespace {

int default_config(int rhs_bits, int n, long long m_vec) {
    // One dword of B per lane and one k-block of gathers in flight per wave:
    // latency is hidden by more resident waves rather than deeper per-wave
    // pipelines.  Up to four waves share a row (one per 16 * 32/bits
    // columns), and when the launch would hold fewer than 2048 waves (8 per
    // CU) the row's k-blocks are also split between two waves.  Measured on
    // MI355X over M = K = 4096, N in {128, 256, 512}, V in {2, 8}, sparsity
    // 0.7-0.98, every pair, and the reference's 2304-deep shapes: within 9 %
    // of the best tile everywhere (the autotuner covers the rest).
    // With eight or more column tiles and 4096 to 8191 waves, eight waves
    // share a row (one workgroup per row at n = 512 for 8-bit B).  Measured
    // on MI355X over M = K = 4096, N = 512, V in {2, 4, 8}, s in {0.7, 0.9,
    // 0.98} for L8-R8 and L16-R8: at V = 8 (4096 waves) within 1 % of the
    // best tile, 1-2 % better than four waves per row at s = 0.7 and 0.98;
    // at V = 2 and 4 (8192 waves and more) four waves per row are better
    // (3-6 % at s = 0.9), and two k-blocks in flight, best there at s <= 0.9,
    // lose 20-30 % at s = 0.98, where the rows are short.
    const int tile_cols = 16 * (32 / rhs_bits);
    const int tiles = (n + tile_cols - 1) / tile_cols;
    if (tiles >= 8) {
        const long long waves8 = m_vec * ((tiles + 7) / 8) * 8;
        if (waves8 >= 4096 && waves8 < 8192) return 13;
    }
    const int tpb = tiles >= 4 ? 4 : (tiles >= 2 ? 2 : 1);
    const long long waves = m_vec * ((tiles + tpb - 1) / tpb) * tpb;
    if (waves < 2048) return tpb == 4 ? 20 : 22;
    return tpb == 4 ? 17 : (tpb == 2 ? 16 : 15);
}

// Four dwords per lane are built for 8- and 16-bit B only (4-bit B would
// hold 32 accumulator tiles per wave); the B-stationary kernel is built for
// 8-bit B.
bool config_supported(int cfg, int rhs_bits) {
    if (kSpmmConfigs[cfg].kind == 1) return rhs_bits == 8;
    return rhs_bits != 4 || kSpmmConfigs[cfg].lane_dwords < 4;
}

// The reference's precision table (SpMM/SpMM/spmm_benchmark.cpp:805-831).
bool pair_supported(int lhs_bits, int rhs_bits) {
    switch (rhs_bits) {
    case 4: return lhs_bits == 4 || lhs_bits == 8 || lhs_bits == 12 || lhs_bits == 16;
    case 8: return lhs_bits == 8 || lhs_bits == 12 || lhs_bits == 16;
    case 16: return lhs_bits == 16;
    default: return false;
    }
}

mc_status spmm_entry(int lhs_bits, int rhs_bits, mc_signedness sign, int config, int m_vec, int vec_length, int n,
                     int k, const int *row_indices, const int *row_offsets, const int *column_indices,
                     const int *values, const int *rhs_matrix, int *output_matrix, hipStream_t stream,
                     int batch = 1, long long values_stride = 0, long long rhs_stride = 0,
                     long long out_stride = 0) {
    if (!pair_supported(lhs_bits, rhs_bits)) return MC_EUNSUPPORTED;
    if (batch < 1 || batch > 65535 || values_stride < 0 || rhs_stride < 0 || out_stride < 0) return MC_EINVAL_SHAPE;
    if ((values_stride | rhs_stride) & 3) return MC_EINVAL_SHAPE;   // every matrix starts dword aligned
    if (sign != MC_SIGNED && sign != MC_UNSIGNED) return MC_EINVAL_ARG;
    if (!valid_v(vec_length)) return MC_EINVAL_V;
    if (m_vec < 0 || n <= 0 || k <= 0) return MC_EINVAL_SHAPE;
    const int cpd = 32 / rhs_bits;
    if (n % cpd != 0) return MC_EINVAL_SHAPE;
    const long long rhs_bytes = (long long)k * n * rhs_bits / 8;
    if (rhs_bytes >= (1LL << 30)) return MC_EINVAL_SHAPE;   // gather offsets stay below 2^30
    if ((long long)m_vec * vec_length * (long long)n >= (1LL << 40)) return MC_EINVAL_SHAPE;
    if (m_vec == 0) return MC_OK;
    if (!row_offsets || !column_indices || !values || !rhs_matrix || !output_matrix) return MC_EINVAL_ARG;
    if (config < -1 || config >= kNumSpmmConfigs) return MC_EINVAL_ARG;
    // a batched launch fills the GPU with (row, tile) x batch workgroups
    int cfg = config >= 0 ? config : default_config(rhs_bits, n, (long long)m_vec * batch);
    if (!config_supported(cfg, rhs_bits)) return MC_EUNSUPPORTED;
    if (batch > 1 && kSpmmConfigs[cfg].kind == 1) return MC_EUNSUPPORTED;   // B-stationary: one matrix
    if (kSpmmConfigs[cfg].kind == 1 && !stationary_fits(cfg, n, k)) {
        if (config >= 0) return MC_EINVAL_SHAPE;
        cfg = 7;   // the gather kernel takes any k
    }
    if ((n / cpd) % kSpmmConfigs[cfg].lane_dwords != 0) {
        if (config >= 0) return MC_EINVAL_SHAPE;
        cfg = 4;   // one dword per lane fits every n
    }

    SpmmArgs a{};
    a.m_vec = m_vec;
    a.n = n;
    a.k = k;
    a.row_indices = row_indices;
    a.row_offsets = row_offsets;
    a.col_indices = column_indices;
    a.values = reinterpret_cast<const unsigned char *>(values);
    a.rhs = reinterpret_cast<const unsigned char *>(rhs_matrix);
    a.out = output_matrix;
    a.rhs_bytes = (unsigned)rhs_bytes;
    a.values_stride = values_stride;
    a.rhs_stride = rhs_stride;
    a.out_stride = out_stride;
    a.batch = batch;
    const bool sgn = sign == MC_SIGNED;
    if (lhs_bits == 8 && rhs_bits == 8) return spmm_launch_l8r8(sgn, vec_length, cfg, a, stream);
    if (lhs_bits == 16 && rhs_bits == 8) return spmm_launch_l16r8(sgn, vec_length, cfg, a, stream);
    if (lhs_bits == 12 && rhs_bits == 8) return spmm_launch_l12r8(sgn, vec_length, cfg, a, stream);
    if (lhs_bits == 16 && rhs_bits == 16) return spmm_launch_l16r16(sgn, vec_length, cfg, a, stream);
    if (lhs_bits == 8 && rhs_bits == 4) return spmm_launch_l8r4(sgn, vec_length, cfg, a, stream);
    if (lhs_bits == 12 && rhs_bits == 4) return spmm_launch_l12r4(sgn, vec_length, cfg, a, stream);
    if (lhs_bits == 16 && rhs_bits == 4) return spmm_launch_l16r4(sgn, vec_length, cfg, a, stream);
    return spmm_launch_l4r4(sgn, vec_length, cfg, a, stream);
}

}  // namespace
}  // namespace mc

extern "C" {

mc_status mc_spmm_l8r8(int m_vec, int vec_length, int n, int k, const int *row_indices, const int *row_offsets,
                       const int *column_indices, const int *values, const int *rhs_matrix, int *output_matrix,
                       mc_signedness sign, hipStream_t stream) {
    return mc::spmm_entry(8, 8, sign, -1, m_vec, vec_length, n, k, row_indices, row_offsets, column_indices, values,
                          rhs_matrix, output_matrix, stream);
}

mc_status mc_spmm_l16r8(int m_vec, int vec_length, int n, int k, const int *row_indices, const int *row_offsets,
                        const int *column_indices, const int *values, const int *rhs_matrix, int *output_matrix,
                        mc_signedness sign, hipStream_t stream) {
    return mc::spmm_entry(16, 8, sign, -1, m_vec, vec_length, n, k, row_indices, row_offsets, column_indices,
                          values, rhs_matrix, output_matrix, stream);
}

mc_status mc_spmm_l8r4(int m_vec, int vec_length, int n, int k, const int *row_indices, const int *row_offsets,
                       const int *column_indices, const int *values, const int *rhs_matrix, int *output_matrix,
                       mc_signedness sign, hipStream_t stream) {
    return mc::spmm_entry(8, 4, sign, -1, m_vec, vec_length, n, k, row_indices, row_offsets, column_indices, values,
                          rhs_matrix, output_matrix, stream);
}

mc_status mc_spmm_l4r4(int m_vec, int vec_length, int n, int k, const int *row_indices, const int *row_offsets,
                       const int *column_indices, const int *values, const int *rhs_matrix, int *output_matrix,
                       mc_signedness sign, hipStream_t stream) {
    return mc::spmm_entry(4, 4, sign, -1, m_vec, vec_length, n, k, row_indices, row_offsets, column_indices, values,
                          rhs_matrix, output_matrix, stream);
}

mc_status mc_spmm_l12r8(int m_vec, int vec_length, int n, int k, const int *row_indices, const int *row_offsets,
                       const int *column_indices, const int *values, const int *rhs_matrix, int *output_matrix,
                       mc_signedness sign, hipStream_t stream) {
    return mc::spmm_entry(12, 8, sign, -1, m_vec, vec_length, n, k, row_indices, row_offsets, column_indices, values,
                          rhs_matrix, output_matrix, stream);
}

mc_status mc_spmm_l16r16(int m_vec, int vec_length, int n, int k, const int *row_indices, const int *row_offsets,
                       const int *column_indices, const int *values, const int *rhs_matrix, int *output_matrix,
                       mc_signedness sign, hipStream_t stream) {
    return mc::spmm_entry(16, 16, sign, -1, m_vec, vec_length, n, k, row_indices, row_offsets, column_indices, values,
                          rhs_matrix, output_matrix, stream);
}

mc_status mc_spmm_l12r4(int m_vec, int vec_length, int n, int k, const int *row_indices, const int *row_offsets,
                       const int *column_indices, const int *values, const int *rhs_matrix, int *output_matrix,
                       mc_signedness sign, hipStream_t stream) {
    return mc::spmm_entry(12, 4, sign, -1, m_vec, vec_length, n, k, row_indices, row_offsets, column_indices, values,
                          rhs_matrix, output_matrix, stream);
}

mc_status mc_spmm_l16r4(int m_vec, int vec_length, int n, int k, const int *row_indices, const int *row_offsets,
                       const int *column_indices, const int *values, const int *rhs_matrix, int *output_matrix,
                       mc_signedness sign, hipStream_t stream) {
    return mc::spmm_entry(16, 4, sign, -1, m_vec, vec_length, n, k, row_indices, row_offsets, column_indices, values,
                          rhs_matrix, output_matrix, stream);
}

int mc_spmm_num_tile_configs(void) { return mc::kNumSpmmConfigs; }

mc_status mc_spmm_get_tile_config(int index, mc_tile_config *out) {
    if (index < 0 || index >= mc::kNumSpmmConfigs || !out) return MC_EINVAL_ARG;
    *out = mc::kSpmmConfigs[index];
    return MC_OK;
}

mc_status mc_spmm_run(int lhs_bits, int rhs_bits, mc_signedness sign, int config, int m_vec, int vec_length, int n,
                      int k, const int *row_indices, const int *row_offsets, const int *column_indices,
                      const int *values, const int *rhs_matrix, int *output_matrix, hipStream_t stream) {
    return mc::spmm_entry(lhs_bits, rhs_bits, sign, config, m_vec, vec_length, n, k, row_indices, row_offsets,
                          column_indices, values, rhs_matrix, output_matrix, stream);
}

mc_status mc_spmm_batched(int lhs_bits, int rhs_bits, mc_signedness sign, int config, int batch, int m_vec,
                          int vec_length, int n, int k, const int *row_indices, const int *row_offsets,
                          const int *column_indices, const int *values, long long values_stride_bytes,
                          const int *rhs_matrix, long long rhs_stride_bytes, int *output_matrix,
                          long long out_stride, hipStream_t stream) {
    return mc::spmm_entry(lhs_bits, rhs_bits, sign, config, m_vec, vec_length, n, k, row_indices, row_offsets,
                          column_indices, values, rhs_matrix, output_matrix, stream, batch, values_stride_bytes,
                          rhs_stride_bytes, out_stride);
}

}  // extern "C"
