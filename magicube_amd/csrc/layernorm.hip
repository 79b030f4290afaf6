// Row LayerNorm on fp16 activations for the encoder benchmark, optionally
// fused with the residual add before it:
//   s = half(x + residual)          (or s = x)
//   y = half((s - mean) * rstd * weight + bias),   rstd = 1 / sqrt(var + eps)
// with mean and the biased variance of s taken in fp32 over the row, as
// torch's nn.LayerNorm computes for fp16 input.  The reference's encoder uses
// torch's LayerNorm and a separate residual add
// (end2end_eval/sparse_transformer_magicube/end_to_end.py:77-89).
//
// Each wave normalises two rows; each lane holds COLS / 64 consecutive
// values of a row, loaded and stored as 8-byte (COLS = 256) or wider vectors.
// The row statistics are two wave reductions (sum, then sum of squared
// deviations): a two-pass variance of values held in registers.
#include <hip/hip_fp16.h>

#include "mc_internal.hpp"

namespace mc {
namespace {

// x + the value of the lane a DPP pattern names (within each row of 16 lanes)
#define LN_DPP_ADD(x, ctrl) \
    ((x) + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), (ctrl), 0xF, 0xF, false)))

// Sum over the wave, the same value in every lane: within each row of 16
// lanes by DPP (lane pairs 1 apart, 2 apart, then the half-row and row
// mirrors; each step adds the same two values in both lanes of a pair),
// then the four row sums read with v_readlane.  No LDS traffic.
__device__ __forceinline__ float ln_wave_sum(float x) {
    x = LN_DPP_ADD(x, 0xB1);     // quad_perm [1, 0, 3, 2]
    x = LN_DPP_ADD(x, 0x4E);     // quad_perm [2, 3, 0, 1]
    x = LN_DPP_ADD(x, 0x141);    // row_half_mirror
    x = LN_DPP_ADD(x, 0x140);    // row_mirror
    const float r0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), 0));
    const float r1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), 16));
    const float r2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), 32));
    const float r3 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), 48));
    return (r0 + r1) + (r2 + r3);
}

// Rows per wave: a wave issues every load of its rows before the first
// reduction (two rows measured fastest, four and one slower).
constexpr int kRowsPerWave = 2;

template <int VPL, int R>   // values per lane: COLS / 64, a multiple of 4; rows per wave
__global__ __launch_bounds__(256) void layernorm_f16_kernel(long long rows, const __half *x, const __half *res,
                                                            const __half *weight, const __half *bias, float eps,
                                                            __half *out) {
    constexpr int COLS = 64 * VPL;
    const int lane = threadIdx.x & 63;
    // rows row0 .. row0 + R - 1 of this wave (wave-uniform bounds)
    const long long row0 = ((long long)blockIdx.x * 4 + (threadIdx.x >> 6)) * R;
    if (row0 >= rows) return;
    const size_t col = (size_t)lane * VPL;
    // every load of the wave's rows goes out before the first reduction
    float v[R][VPL];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (row0 + r >= rows) break;
        const size_t base = (size_t)(row0 + r) * COLS + col;
#pragma unroll
        for (int q = 0; q < VPL; q += 4) {
            const __half2 *px = reinterpret_cast<const __half2 *>(x + base + q);
            __half2 a0 = px[0], a1 = px[1];
            if (res) {
                const __half2 *pr = reinterpret_cast<const __half2 *>(res + base + q);
                a0 = __hadd2(a0, pr[0]);     // rounded to fp16, as a separate add would store it
                a1 = __hadd2(a1, pr[1]);
            }
            v[r][q] = __low2float(a0);
            v[r][q + 1] = __high2float(a0);
            v[r][q + 2] = __low2float(a1);
            v[r][q + 3] = __high2float(a1);
        }
    }
    float wv[VPL], bv[VPL];
#pragma unroll
    for (int q = 0; q < VPL; ++q) {
        wv[q] = weight ? __half2float(weight[col + q]) : 1.f;
        bv[q] = bias ? __half2float(bias[col + q]) : 0.f;
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (row0 + r >= rows) break;
        float sum = 0.f;
#pragma unroll
        for (int q = 0; q < VPL; ++q) sum += v[r][q];
        const float mean = ln_wave_sum(sum) * (1.0f / COLS);
        float sq = 0.f;
#pragma unroll
        for (int q = 0; q < VPL; ++q) {
            const float d = v[r][q] - mean;
            sq += d * d;
        }
        const float rstd = rsqrtf(ln_wave_sum(sq) * (1.0f / COLS) + eps);
        const size_t base = (size_t)(row0 + r) * COLS + col;
#pragma unroll
        for (int q = 0; q < VPL; q += 2) {
            const float y0 = (v[r][q] - mean) * rstd * wv[q] + bv[q];
            const float y1 = (v[r][q + 1] - mean) * rstd * wv[q + 1] + bv[q + 1];
            *reinterpret_cast<__half2 *>(out + base + q) = __floats2half2_rn(y0, y1);
        }
    }
}

}  // namespace
}  // namespace mc

extern "C" {

mc_status mc_layernorm_f16(long long rows, int cols, const void *x, const void *residual, const void *weight,
                           const void *bias, float eps, void *out, hipStream_t stream) {
    if (rows < 0 || (cols != 256 && cols != 512 && cols != 1024)) return MC_EINVAL_SHAPE;
    if (rows == 0) return MC_OK;
    if (!x || !out) return MC_EINVAL_ARG;
    // 8-byte vector loads of four halves
    const auto aligned = [](const void *p) { return (reinterpret_cast<uintptr_t>(p) & 7u) == 0; };
    if (!aligned(x) || !aligned(out) || (residual && !aligned(residual))) return MC_EINVAL_ARG;
    constexpr int R = mc::kRowsPerWave;
    if ((rows + 4 * R - 1) / (4 * R) > 0x7FFFFFFFLL) return MC_EINVAL_SHAPE;
    const dim3 grid((unsigned)((rows + 4 * R - 1) / (4 * R))), block(256);
    const auto *px = static_cast<const __half *>(x);
    const auto *pr = static_cast<const __half *>(residual);
    const auto *pw = static_cast<const __half *>(weight);
    const auto *pb = static_cast<const __half *>(bias);
    auto *po = static_cast<__half *>(out);
    switch (cols) {
    case 256: hipLaunchKernelGGL((mc::layernorm_f16_kernel<4, R>), grid, block, 0, stream, rows, px, pr, pw, pb, eps, po); break;
    case 512: hipLaunchKernelGGL((mc::layernorm_f16_kernel<8, R>), grid, block, 0, stream, rows, px, pr, pw, pb, eps, po); break;
    default: hipLaunchKernelGGL((mc::layernorm_f16_kernel<16, R>), grid, block, 0, stream, rows, px, pr, pw, pb, eps, po); break;
    }
    return mc::check_launch();
}

}  // extern "C"
