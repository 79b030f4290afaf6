// C ABI of the SDDMM family: validation and dispatch.
#include "mc_internal.hpp"
#include "sddmm_kernel.hpp"

namespace mc {
namespace {

// One workgroup per vector row.  A single 64-deep chunk (K = 64) gives each
// of 16 waves one 16-slot tile at a time, so a row's gathers are spread over
// more waves; deeper K keeps 8 waves with up to four tiles each, which hold
// more of the row's A operands per wave.  Measured on MI355X at
// M = N = 4096, V = 8: K = 64 5 % faster with 16 waves, K = 128 and 512
// 7-11 % slower.
template <int V, int BITS, bool SGN, int NKC, bool REF>
mc_status sddmm_launch(const SddmmArgs &a, int batch, hipStream_t s) {
    constexpr int W = NKC == 1 ? 16 : 8;
    constexpr int TPW = NKC == 1 ? 1 : (NKC >= 4 ? 1 : 4 / NKC);
    hipLaunchKernelGGL((sddmm_kernel<V, BITS, SGN, NKC, W, TPW, REF>), dim3((unsigned)a.m_vec, (unsigned)batch),
                       dim3(64 * W), 0, s, a);
    return check_launch();
}

template <int V, int BITS, bool SGN, bool REF>
mc_status sddmm_launch_k(int nkc, const SddmmArgs &a, int batch, hipStream_t s) {
    switch (nkc) {
    case 1: return sddmm_launch<V, BITS, SGN, 1, REF>(a, batch, s);
    case 2: return sddmm_launch<V, BITS, SGN, 2, REF>(a, batch, s);
    case 4: return sddmm_launch<V, BITS, SGN, 4, REF>(a, batch, s);
    case 8: return sddmm_launch<V, BITS, SGN, 8, REF>(a, batch, s);
    default: return MC_EUNSUPPORTED;
    }
}

template <int BITS, bool SGN, bool REF>
mc_status sddmm_launch_v(int V, int nkc, const SddmmArgs &a, int batch, hipStream_t s) {
    switch (V) {
    case 2: return sddmm_launch_k<2, BITS, SGN, REF>(nkc, a, batch, s);
    case 4: return sddmm_launch_k<4, BITS, SGN, REF>(nkc, a, batch, s);
    case 8: return sddmm_launch_k<8, BITS, SGN, REF>(nkc, a, batch, s);
    default: return MC_EINVAL_V;
    }
}

template <int BITS>
mc_status sddmm_launch_all(bool sgn, bool ref, int V, int nkc, const SddmmArgs &a, int batch, hipStream_t s) {
    if (sgn)
        return ref ? sddmm_launch_v<BITS, true, true>(V, nkc, a, batch, s)
                   : sddmm_launch_v<BITS, true, false>(V, nkc, a, batch, s);
    return ref ? sddmm_launch_v<BITS, false, true>(V, nkc, a, batch, s)
               : sddmm_launch_v<BITS, false, false>(V, nkc, a, batch, s);
}

mc_status sddmm_entry(int bits, mc_signedness sign, int layout, int m_vec, int k, int n, const int *row_indices,
                      const int *row_offsets, const int *col_indices, const int *lhs_matrix, const int *rhs_matrix,
                      int *output_values, int vec_length, hipStream_t stream, int batch = 1,
                      long long lhs_stride = 0, long long rhs_stride = 0, long long out_stride = 0) {
    if (batch < 1 || batch > 65535 || lhs_stride < 0 || rhs_stride < 0 || out_stride < 0) return MC_EINVAL_SHAPE;
    if ((lhs_stride | rhs_stride) & 3) return MC_EINVAL_SHAPE;   // every matrix starts dword aligned
    if (bits != 16 && bits != 8 && bits != 4) return MC_EUNSUPPORTED;
    if (sign != MC_SIGNED && sign != MC_UNSIGNED) return MC_EINVAL_ARG;
    if (layout != MC_SDDMM_OUT_REFERENCE && layout != MC_SDDMM_OUT_VECTOR) return MC_EINVAL_ARG;
    if (!valid_v(vec_length)) return MC_EINVAL_V;
    if (m_vec < 0 || n <= 0 || k <= 0 || k % 64 != 0) return MC_EINVAL_SHAPE;
    // 64-deep chunks held per pass: the smallest of 1, 2, 4, 8 that covers k,
    // else 8 with several passes
    const int chunks = k / 64;
    const int nkc = chunks <= 1 ? 1 : chunks <= 2 ? 2 : chunks <= 4 ? 4 : 8;
    const long long lhs_bytes = (long long)m_vec * vec_length * k * bits / 8;
    const long long rhs_bytes = (long long)n * k * bits / 8;
    if (lhs_bytes >= (1LL << 30) || rhs_bytes >= (1LL << 30)) return MC_EINVAL_SHAPE;
    if (m_vec == 0) return MC_OK;
    if (!row_offsets || !col_indices || !lhs_matrix || !rhs_matrix || !output_values) return MC_EINVAL_ARG;
    SddmmArgs a{};
    a.m_vec = m_vec;
    a.k = k;
    a.n = n;
    a.row_indices = row_indices;
    a.row_offsets = row_offsets;
    a.col_indices = col_indices;
    a.lhs = reinterpret_cast<const unsigned char *>(lhs_matrix);
    a.rhs = reinterpret_cast<const unsigned char *>(rhs_matrix);
    a.out = output_values;
    a.lhs_bytes = (unsigned)lhs_bytes;
    a.rhs_bytes = (unsigned)rhs_bytes;
    a.lhs_stride = lhs_stride;
    a.rhs_stride = rhs_stride;
    a.out_stride = out_stride;
    const bool sgn = sign == MC_SIGNED, ref = layout == MC_SDDMM_OUT_REFERENCE;
    if (bits == 16) return sddmm_launch_all<16>(sgn, ref, vec_length, nkc, a, batch, stream);
    return bits == 8 ? sddmm_launch_all<8>(sgn, ref, vec_length, nkc, a, batch, stream)
                     : sddmm_launch_all<4>(sgn, ref, vec_length, nkc, a, batch, stream);
}

}  // namespace
}  // namespace mc

extern "C" {

mc_status mc_sddmm_l8r8(int m_vec, int k, int n, const int *row_indices, const int *row_offsets,
                        const int *col_indices, const int *lhs_matrix, const int *rhs_matrix, int *output_values,
                        int vec_length, mc_signedness sign, hipStream_t stream) {
    return mc::sddmm_entry(8, sign, MC_SDDMM_OUT_REFERENCE, m_vec, k, n, row_indices, row_offsets, col_indices,
                           lhs_matrix, rhs_matrix, output_values, vec_length, stream);
}

mc_status mc_sddmm_l4r4(int m_vec, int k, int n, const int *row_indices, const int *row_offsets,
                        const int *col_indices, const int *lhs_matrix, const int *rhs_matrix, int *output_values,
                        int vec_length, mc_signedness sign, hipStream_t stream) {
    return mc::sddmm_entry(4, sign, MC_SDDMM_OUT_REFERENCE, m_vec, k, n, row_indices, row_offsets, col_indices,
                           lhs_matrix, rhs_matrix, output_values, vec_length, stream);
}

mc_status mc_sddmm_l16r16(int m_vec, int k, int n, const int *row_indices, const int *row_offsets,
                          const int *col_indices, const int *lhs_matrix, const int *rhs_matrix, int *output_values,
                          int vec_length, mc_signedness sign, hipStream_t stream) {
    return mc::sddmm_entry(16, sign, MC_SDDMM_OUT_REFERENCE, m_vec, k, n, row_indices, row_offsets, col_indices,
                           lhs_matrix, rhs_matrix, output_values, vec_length, stream);
}

mc_status mc_sddmm_run(int bits, mc_signedness sign, int layout, int m_vec, int k, int n, const int *row_indices,
                       const int *row_offsets, const int *col_indices, const int *lhs_matrix,
                       const int *rhs_matrix, int *output_values, int vec_length, hipStream_t stream) {
    return mc::sddmm_entry(bits, sign, layout, m_vec, k, n, row_indices, row_offsets, col_indices, lhs_matrix,
                           rhs_matrix, output_values, vec_length, stream);
}

mc_status mc_sddmm_batched(int bits, mc_signedness sign, int layout, int batch, int m_vec, int k, int n,
                           const int *row_indices, const int *row_offsets, const int *col_indices,
                           const int *lhs_matrix, long long lhs_stride_bytes, const int *rhs_matrix,
                           long long rhs_stride_bytes, int *output_values, long long out_stride, int vec_length,
                           hipStream_t stream) {
    return mc::sddmm_entry(bits, sign, layout, m_vec, k, n, row_indices, row_offsets, col_indices, lhs_matrix,
                           rhs_matrix, output_values, vec_length, stream, batch, lhs_stride_bytes,
                           rhs_stride_bytes, out_stride);
}

}  // extern "C"
