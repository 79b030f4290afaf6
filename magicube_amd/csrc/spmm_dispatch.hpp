// Template dispatch for the SpMM kernel family: (V, bits, signedness, tile
// config) -> one instantiation.  Each pair is instantiated in its own
// translation unit (spmm_l*.hip) so the build parallelizes.
#pragma once

#include "mc_internal.hpp"
#include "spmm_kernel.hpp"
#include "spmm_stationary.hpp"

namespace mc {

// Tile configurations {lane_dwords, tiles_per_block, split_k, depth, kind,
// waves_per_simd}: 0-7 and 12-23 gather kernel (four dwords per lane: 8- and
// 16-bit B); 8-11 B-stationary kernel (8-bit B), for which the fields are {B dwords per slice row, waves per workgroup,
// vector rows per workgroup, prefetch depth}.  waves_per_simd > 1 makes the
// register allocation fit that many waves per SIMD.  13 and 14 are config 7
// with one and two k-blocks of gathers in flight per wave instead of three
// (the software-pipelining ablation).  15-23 vary the waves per row (1, 2,
// 4, 8), the split of a row's k-blocks over waves and the depth; the defaults
// (15, 16, 17, 20, 22) come from them.
constexpr int kNumSpmmConfigs = 24;
constexpr mc_tile_config kSpmmConfigs[kNumSpmmConfigs] = {
    {2, 4, 1, 2, 0, 0}, {2, 4, 2, 2, 0, 0}, {4, 2, 2, 2, 0, 0}, {4, 2, 4, 2, 0, 0},
    {1, 4, 2, 3, 0, 0}, {2, 2, 2, 2, 0, 0}, {4, 1, 4, 2, 0, 0}, {1, 8, 1, 3, 0, 0},
    {8, 16, 32, 4, 1, 0}, {8, 8, 32, 4, 1, 0}, {16, 16, 32, 4, 1, 0}, {8, 16, 16, 4, 1, 0},
    {1, 4, 1, 2, 0, 6}, {1, 8, 1, 1, 0, 0}, {1, 8, 1, 2, 0, 0},
    {1, 1, 1, 1, 0, 0}, {1, 2, 1, 1, 0, 0}, {1, 4, 1, 1, 0, 0}, {1, 2, 1, 2, 0, 0},
    {1, 4, 1, 3, 0, 0}, {1, 4, 2, 1, 0, 0}, {1, 8, 2, 1, 0, 0}, {1, 2, 2, 1, 0, 0},
    {1, 4, 2, 2, 0, 0},
};

// Whether a B-stationary config fits the shape (slice width divides n, the
// K rows of the slice fit in LDS).
inline bool stationary_fits(int cfg, int n, int k) {
    const mc_tile_config &c = kSpmmConfigs[cfg];
    const int tb = 4 * c.lane_dwords;
    return c.kind == 1 && n % tb == 0 && (size_t)(k + 1) * tb <= (size_t)kStatLdsBytes;
}

template <int V, int AB, int BB, bool SGN, int LW, int TPB, int SK, int D, int WPS = 1>
mc_status spmm_launch_one(SpmmArgs a, hipStream_t stream) {
    constexpr int TILE_N = 16 * LW * (32 / BB);
    if ((a.n / (32 / BB)) % LW != 0) return MC_EUNSUPPORTED;
    a.tiles_n = (a.n + TILE_N - 1) / TILE_N;
    const int groups = (a.tiles_n + TPB - 1) / TPB;
    const long long blocks = (long long)a.m_vec * groups;
    if (blocks <= 0) return MC_OK;
    if (blocks > 0x7fffffffLL) return MC_EINVAL_SHAPE;
    hipLaunchKernelGGL((spmm_kernel<V, AB, BB, SGN, LW, TPB, SK, D, WPS>), dim3((unsigned)blocks, (unsigned)a.batch),
                       dim3(64 * TPB * SK), 0, stream, a);
    return check_launch();
}

template <int V, int AB, int BB, bool SGN>
mc_status spmm_launch_cfg(int cfg, const SpmmArgs &a, hipStream_t s) {
    switch (cfg) {
    case 0: return spmm_launch_one<V, AB, BB, SGN, 2, 4, 1, 2>(a, s);
    case 1: return spmm_launch_one<V, AB, BB, SGN, 2, 4, 2, 2>(a, s);
    case 2:
        if constexpr (BB != 4) return spmm_launch_one<V, AB, BB, SGN, 4, 2, 2, 2>(a, s);
        else return MC_EUNSUPPORTED;
    case 3:
        if constexpr (BB != 4) return spmm_launch_one<V, AB, BB, SGN, 4, 2, 4, 2>(a, s);
        else return MC_EUNSUPPORTED;
    case 4: return spmm_launch_one<V, AB, BB, SGN, 1, 4, 2, 3>(a, s);
    case 5: return spmm_launch_one<V, AB, BB, SGN, 2, 2, 2, 2>(a, s);
    case 6:
        if constexpr (BB != 4) return spmm_launch_one<V, AB, BB, SGN, 4, 1, 4, 2>(a, s);
        else return MC_EUNSUPPORTED;
    case 7: return spmm_launch_one<V, AB, BB, SGN, 1, 8, 1, 3>(a, s);
    case 8:
        if constexpr (BB == 8) return spmm_stationary_launch<V, AB, SGN, 32, 16, 4>(a, 32, s);
        else return MC_EUNSUPPORTED;
    case 9:
        if constexpr (BB == 8) return spmm_stationary_launch<V, AB, SGN, 32, 8, 4>(a, 32, s);
        else return MC_EUNSUPPORTED;
    case 10:
        if constexpr (BB == 8) return spmm_stationary_launch<V, AB, SGN, 64, 16, 4>(a, 32, s);
        else return MC_EUNSUPPORTED;
    case 11:
        if constexpr (BB == 8) return spmm_stationary_launch<V, AB, SGN, 32, 16, 4>(a, 16, s);
        else return MC_EUNSUPPORTED;
    case 12: return spmm_launch_one<V, AB, BB, SGN, 1, 4, 1, 2, 6>(a, s);
    case 13: return spmm_launch_one<V, AB, BB, SGN, 1, 8, 1, 1>(a, s);
    case 14: return spmm_launch_one<V, AB, BB, SGN, 1, 8, 1, 2>(a, s);
    case 15: return spmm_launch_one<V, AB, BB, SGN, 1, 1, 1, 1>(a, s);
    case 16: return spmm_launch_one<V, AB, BB, SGN, 1, 2, 1, 1>(a, s);
    case 17: return spmm_launch_one<V, AB, BB, SGN, 1, 4, 1, 1>(a, s);
    case 18: return spmm_launch_one<V, AB, BB, SGN, 1, 2, 1, 2>(a, s);
    case 19: return spmm_launch_one<V, AB, BB, SGN, 1, 4, 1, 3>(a, s);
    case 20: return spmm_launch_one<V, AB, BB, SGN, 1, 4, 2, 1>(a, s);
    case 21: return spmm_launch_one<V, AB, BB, SGN, 1, 8, 2, 1>(a, s);
    case 22: return spmm_launch_one<V, AB, BB, SGN, 1, 2, 2, 1>(a, s);
    case 23: return spmm_launch_one<V, AB, BB, SGN, 1, 4, 2, 2>(a, s);
    default: return MC_EINVAL_ARG;
    }
}

template <int AB, int BB>
mc_status spmm_launch_pair(bool sgn, int V, int cfg, const SpmmArgs &a, hipStream_t s) {
    if (sgn) {
        switch (V) {
        case 2: return spmm_launch_cfg<2, AB, BB, true>(cfg, a, s);
        case 4: return spmm_launch_cfg<4, AB, BB, true>(cfg, a, s);
        case 8: return spmm_launch_cfg<8, AB, BB, true>(cfg, a, s);
        }
    } else {
        switch (V) {
        case 2: return spmm_launch_cfg<2, AB, BB, false>(cfg, a, s);
        case 4: return spmm_launch_cfg<4, AB, BB, false>(cfg, a, s);
        case 8: return spmm_launch_cfg<8, AB, BB, false>(cfg, a, s);
        }
    }
    return MC_EINVAL_V;
}

mc_status spmm_launch_l8r8(bool sgn, int V, int cfg, const SpmmArgs &a, hipStream_t s);
mc_status spmm_launch_l16r8(bool sgn, int V, int cfg, const SpmmArgs &a, hipStream_t s);
mc_status spmm_launch_l8r4(bool sgn, int V, int cfg, const SpmmArgs &a, hipStream_t s);
mc_status spmm_launch_l4r4(bool sgn, int V, int cfg, const SpmmArgs &a, hipStream_t s);
mc_status spmm_launch_l12r8(bool sgn, int V, int cfg, const SpmmArgs &a, hipStream_t s);
mc_status spmm_launch_l12r4(bool sgn, int V, int cfg, const SpmmArgs &a, hipStream_t s);
mc_status spmm_launch_l16r4(bool sgn, int V, int cfg, const SpmmArgs &a, hipStream_t s);
mc_status spmm_launch_l16r16(bool sgn, int V, int cfg, const SpmmArgs &a, hipStream_t s);

}  // namespace mc
