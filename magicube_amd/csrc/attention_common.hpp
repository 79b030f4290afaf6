// Quantized sparse-attention building blocks shared by the fused batched
// kernel and the one-op-per-launch kernels, so both produce identical
// integers (int32 logits, int8 probabilities, int32 outputs).
//
// Layer semantics (end2end_eval/sparse_transformer_magicube/spattention.py:
// 182-216): per (sample, head) matrix with sequence length L and head dim 64,
//   S = Qq Kq^T at the vector-sparse mask          (int32, exact)
//   s = float(S) * alpha        alpha = 1 / (scale_q * scale_k * sqrt(64))
//   reference requantization (q_softmax_kernel.cu:132-147):
//     P = trunc(min(exp(s - max_row) / sum_row * scale_p, 127))   (int8 >= 0)
//     O = float(P Vq) * o_scale          o_scale = 1 / (scale_v * scale_p)
//   row-scaled requantization (full int8 range in every row, rounded):
//     P = rint(min(exp(s - max_row) * scale_p, 127))
//     O = (float(P Vq) * o_scale) * (1 / sum_row)
// Rows are the V matrix rows of one vector row; max/sum run over the row's
// nonzero slots.
//
// Canonical softmax statistics: a row's slots are visited in chunks of 64
// (four 16-slot MFMA tiles, slot order).  Per chunk: exact max M_c; running
// max m' = max(m, M_c); lane c first adds its four slots c, 16 + c, 32 + c,
// 48 + c as (e0 + e1) + (e2 + e3), then the 16 lane sums are added as a
// pairwise tree over lane order; running sum l = l * exp(m - m') + chunk sum.  Both kernels evaluate these statistics
// on the same register layout (MFMA C layout: lane (c, g) holds rows 4g+i of
// slot 16t + c) with the same instructions.  The fused kernel first moves
// the V real rows out of the MFMA tile's padding lanes (rows are dealt to
// the four 16-lane groups instead of four rows per group), so it keeps fewer
// rows per lane; each row is still held by the 16 lanes c of one group with
// slots 16t + c, and its statistics take the same operations in the same
// order.
#pragma once

#include "mc_device.hpp"

namespace mc {

constexpr int kHeadDim = 64;

// Butterfly reductions within each 16-lane row of the wave, on DPP lane
// moves (no LDS round trip): swap neighbours, swap pairs (quad_perm), then
// combine the two quads of each 8-lane half (row_half_mirror) and the two
// halves (row_mirror).  Each step adds the same two partial results in every
// lane that holds them (in commuted order), so every lane of the row ends
// with the same bits, equal to the xor butterfly over lane distances 1, 2,
// 4, 8: ((x0+x1)+(x2+x3)) + ((x4+x5)+(x6+x7)) ... for a sum.
template <int CTRL> __device__ __forceinline__ float dpp_move(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, 0xF, 0xF, false));
}
constexpr int kDppSwap1 = 0xB1;        // quad_perm [1, 0, 3, 2]
constexpr int kDppSwap2 = 0x4E;        // quad_perm [2, 3, 0, 1]
constexpr int kDppHalfMirror = 0x141;  // lane i <- lane 7 - i within 8
constexpr int kDppMirror = 0x140;      // lane i <- lane 15 - i within 16
__device__ __forceinline__ float row16_max(float x) {
    x = fmaxf(x, dpp_move<kDppSwap1>(x));
    x = fmaxf(x, dpp_move<kDppSwap2>(x));
    x = fmaxf(x, dpp_move<kDppHalfMirror>(x));
    x = fmaxf(x, dpp_move<kDppMirror>(x));
    return x;
}
__device__ __forceinline__ float row16_sum(float x) {
    x = x + dpp_move<kDppSwap1>(x);
    x = x + dpp_move<kDppSwap2>(x);
    x = x + dpp_move<kDppHalfMirror>(x);
    x = x + dpp_move<kDppMirror>(x);
    return x;
}

// Online statistics of one chunk for the HOME rows held by this lane.
// s[t][i]: scaled logit of slot 16t + c for the lane's row i (-inf when masked).
template <int HOME, int W>
__device__ __forceinline__ void softmax_chunk_update(const float (&s)[4][W], float (&m)[HOME], float (&l)[HOME]) {
    static_assert(HOME <= W, "rows held per lane");
#pragma unroll
    for (int i = 0; i < HOME; ++i) {
        float cm = fmaxf(fmaxf(s[0][i], s[1][i]), fmaxf(s[2][i], s[3][i]));
        cm = row16_max(cm);
        const float mn = fmaxf(m[i], cm);        // -inf only while every slot so far is masked
        // masked slots give exp(-inf) = 0 exactly; the offset is 0 rather
        // than -inf while the row has no slot yet (-inf - -inf would be NaN)
        const float mz = mn == -INFINITY ? 0.f : mn;
        float e[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) e[t] = __expf(s[t][i] - mz);
        const float csum = row16_sum((e[0] + e[1]) + (e[2] + e[3]));
        const float scale_old = (m[i] == -INFINITY) ? 0.f : __expf(m[i] - mn);
        l[i] = l[i] * scale_old + csum;
        m[i] = mn;
    }
}

constexpr int kPModeReference = 0;
constexpr int kPModeRowScaled = 1;
constexpr int kPModeReference4 = 2;   // reference requantization into 4-bit P

// Row max as quantize_prob takes it: 0 for a row without slots.
__device__ __forceinline__ float finite_row_max(float m) { return m == -INFINITY ? 0.f : m; }

// Quantized probability of one logit given the final row statistics
// (m from finite_row_max); a masked slot (s = -inf) gives exp(-inf) = 0.
// PM fixed at compile time (the fused kernel's chunk loops), or -1 for a
// run-time p_mode; both evaluate the same expressions.
template <int PM = -1>
__device__ __forceinline__ int quantize_prob(float s, float m, float inv_l, float p_scale, int p_mode) {
    const int mode = PM >= 0 ? PM : p_mode;
    const float e = __expf(s - m);
    if (mode != kPModeRowScaled)   // truncation (p >= 0)
        return (int)fminf(e * inv_l * p_scale, mode == kPModeReference4 ? 7.f : 127.f);
    return (int)rintf(fminf(e * p_scale, 127.f));
}


// Moves P from the lanes' row layout (lane (c, g): rows row0 + i of slot
// 16t + c; row0 = 4g in the MFMA C layout) to the MFMA A layout (lane (r, g):
// row r, slots 16g..16g+15) through a per-wave 16 x 64 byte LDS tile laid
// out [row][slot].
template <int HOME, int W>
__device__ __forceinline__ v4i p_to_a_layout(unsigned char *tile, const int (&pq)[4][W], int lane, bool home,
                                             int row0) {
    static_assert(HOME <= W, "rows held per lane");
    const int c = lane & 15, g = lane >> 4;
    if (home) {
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int i = 0; i < HOME; ++i) tile[(row0 + i) * 64 + 16 * t + c] = (unsigned char)pq[t][i];
    }
    // The tile is private to the wave and LDS operations of one wave complete
    // in order, so wavefront-scope fences (no waits) are enough: they keep
    // the compiler from moving the read above the writes (and the next writes
    // above this read).  Workgroup scope would also wait for every gather in
    // flight that lands in LDS.  Rows >= V are never written and only feed
    // output rows that are discarded.
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const v4i a = *reinterpret_cast<const v4i *>(tile + c * 64 + 16 * g);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    return a;
}

}  // namespace mc
