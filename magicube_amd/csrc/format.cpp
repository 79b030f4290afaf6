// Host-side vector-CSR preparation: the reference layouts built slot by slot.
//
// For a pair with row alignment MMA_K (16, or 32 with a 4-bit operand), slot
// s of the padded arrays belongs to group G = s / MMA_K at position
// u = s % MMA_K.  Per group the prepared values hold, for each vector lane v,
// a 16-byte row at offset 16*v of each plane:
//   L8-R8   1 byte plane  [v][16 bytes]                (16*V bytes/group)
//   L16-R8  low-byte plane, then high-byte plane        (32*V bytes/group)
//   L12-R8  as L16-R8 with the high byte masked to bits 8..11
//   L16-R16 as L16-R8
//   L8-R4   low-nibble plane, then high-nibble plane; slot u at byte u/2,
//           nibble u%2                                  (32*V bytes/group)
//   L16-R4  nibble planes q = 0..3 (bits 4q..4q+3) in that order, slots as
//           L8-R4                                       (64*V bytes/group)
//   L12-R4  as L16-R4 with plane 3 left zero            (64*V bytes/group)
//   L4-R4   1 nibble plane, slot u at byte u/2, nibble u%2 (16*V bytes/group)
// which is the byte image BmFN produces (SpMM/SpMM/spmm_benchmark.cpp:246-377).
// 12-bit lanes arrive in 16-bit containers.
#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "../../include/magicube_hip.h"

namespace {

inline unsigned lane_bits(const unsigned char *vec, int lane, int bits) {
    const int bit = lane * bits;
    unsigned v = 0;
    for (int b = 0; b < bits; b += 4) {            // bits is 4, 8 or 16 (container width)
        const int q = bit + b;
        v |= (unsigned)((vec[q >> 3] >> (q & 7)) & 0xF) << b;
    }
    return v;
}

}  // namespace

extern "C" {

int mc_format_row_alignment(int lhs_bits, int rhs_bits) { return (lhs_bits == 4 || rhs_bits == 4) ? 32 : 16; }

int mc_format_align_rows(int m_vec, const int *row_offsets, int alignment, int *ap) {
    if (m_vec < 0 || alignment <= 0 || (m_vec > 0 && (!row_offsets || !ap))) return -1;
    long long total = 0;
    for (int i = 0; i < m_vec; ++i) {
        const int nnz = row_offsets[i + 1] - row_offsets[i];
        if (nnz < 0) return -1;
        ap[2 * i] = (int)total;
        ap[2 * i + 1] = (int)(total + nnz);
        total += (long long)(nnz + alignment - 1) / alignment * alignment;
        if (total > 0x7fffffffLL) return -1;
    }
    return (int)total;
}

mc_status mc_format_align_cols(int m_vec, const int *row_offsets, const int *col_indices, const int *ap,
                               int aligned_num, int shuffle, int *out) {
    if (m_vec < 0 || aligned_num < 0 || (aligned_num > 0 && !out)) return MC_EINVAL_ARG;
    std::fill(out, out + aligned_num, -1);
    for (int i = 0; i < m_vec; ++i) {
        const int b = row_offsets[i], e = row_offsets[i + 1];
        if (ap[2 * i] + (e - b) > aligned_num) return MC_EINVAL_SHAPE;
        std::memcpy(out + ap[2 * i], col_indices + b, sizeof(int) * (size_t)(e - b));
    }
    if (shuffle) {
        // within each group of 8 slots: even positions first, then odd
        int tmp[8];
        for (int g = 0; g + 8 <= aligned_num; g += 8) {
            for (int j = 0; j < 8; ++j) tmp[(j & 1) * 4 + (j >> 1)] = out[g + j];
            std::memcpy(out + g, tmp, sizeof(tmp));
        }
    }
    return MC_OK;
}

mc_status mc_format_prepare_values(int lhs_bits, int rhs_bits, int V, int m_vec, const int *row_offsets,
                                   const int *ap, int aligned_num, const void *values, void *out_) {
    const bool ok_pair = rhs_bits == 4   ? (lhs_bits == 4 || lhs_bits == 8 || lhs_bits == 12 || lhs_bits == 16)
                         : rhs_bits == 8 ? (lhs_bits == 8 || lhs_bits == 12 || lhs_bits == 16)
                                         : (rhs_bits == 16 && lhs_bits == 16);
    if (!ok_pair) return MC_EUNSUPPORTED;
    if (V != 2 && V != 4 && V != 8) return MC_EINVAL_V;
    if (m_vec < 0 || aligned_num < 0) return MC_EINVAL_SHAPE;
    const int mma_k = mc_format_row_alignment(lhs_bits, rhs_bits);
    if (aligned_num % mma_k != 0) return MC_EINVAL_SHAPE;
    const int cbits = lhs_bits == 12 ? 16 : lhs_bits;   // container width
    const size_t vec_bytes = (size_t)V * cbits / 8;
    const size_t total = (size_t)aligned_num * vec_bytes;
    unsigned char *out = static_cast<unsigned char *>(out_);
    const unsigned char *vals = static_cast<const unsigned char *>(values);
    std::memset(out, 0, total);
    const size_t group_bytes = (size_t)mma_k * vec_bytes;
    for (int i = 0; i < m_vec; ++i) {
        const int b = row_offsets[i], e = row_offsets[i + 1];
        for (int j = b; j < e; ++j) {
            const size_t s = (size_t)ap[2 * i] + (j - b);
            const size_t G = s / mma_k;
            const int u = (int)(s % mma_k);
            unsigned char *grp = out + G * group_bytes;
            const unsigned char *vec = vals + (size_t)j * vec_bytes;
            for (int v = 0; v < V; ++v) {
                const unsigned a = lane_bits(vec, v, cbits) & ((1u << lhs_bits) - 1u);
                if (lhs_bits == 8 && rhs_bits == 8) {
                    grp[16 * v + u] = (unsigned char)a;
                } else if (lhs_bits >= 12 && rhs_bits != 4) {
                    grp[16 * v + u] = (unsigned char)(a & 0xFF);
                    grp[16 * V + 16 * v + u] = (unsigned char)(a >> 8);
                } else if (lhs_bits >= 12) {  // 4-bit B: nibble planes at 16*V*q
                    for (int q = 0; q < lhs_bits / 4; ++q)
                        grp[16 * V * q + 16 * v + u / 2] |= (unsigned char)(((a >> (4 * q)) & 0xF) << (4 * (u & 1)));
                } else if (lhs_bits == 8) {  // 4-bit B: nibble planes
                    grp[16 * v + u / 2] |= (unsigned char)((a & 0xF) << (4 * (u & 1)));
                    grp[16 * V + 16 * v + u / 2] |= (unsigned char)((a >> 4) << (4 * (u & 1)));
                } else {
                    grp[16 * v + u / 2] |= (unsigned char)((a & 0xF) << (4 * (u & 1)));
                }
            }
        }
    }
    return MC_OK;
}

mc_status mc_format_sorted_rows(int m_vec, const int *row_offsets, int *row_indices) {
    if (m_vec < 0 || (m_vec > 0 && (!row_offsets || !row_indices))) return MC_EINVAL_ARG;
    std::iota(row_indices, row_indices + m_vec, 0);
    std::stable_sort(row_indices, row_indices + m_vec, [row_offsets](int a, int b) {
        return row_offsets[a + 1] - row_offsets[a] > row_offsets[b + 1] - row_offsets[b];
    });
    return MC_OK;
}

}  // extern "C"
