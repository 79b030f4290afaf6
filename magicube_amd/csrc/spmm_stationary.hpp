// B-stationary SpMM for 8-bit B (pairs L8-R8 and L16-R8) on gfx950.
//
// A workgroup owns a TB-byte column slice of B (TB = 32 or 64 columns) and a
// group of vector rows.  It copies that slice of all K rows of B, plus one
// zero row, into LDS once, then every wave serves its vector rows from LDS:
//
//   * per 64-slot k-block a lane holds, for each of two 8-slot halves of its
//     lane group's 16 slots, the LDS address of (row = column index of slot
//     s0 + r/2, 8-column segment r%2); one ds_read_b64_tr_b8 then hands lane
//     c of the group the bytes of column c for those 8 slots, which is the
//     MFMA B operand layout (16 k-bytes of one column per lane) with no
//     register transposes;
//   * padding slots (index -1) and indices outside [0, K) address the zero row;
//   * the A operand, the digit planes, the unsigned-B offset and the epilogue
//     are those of the gather kernel (spmm_kernel.hpp).
//
// A wave's k-blocks over all of its rows form one sequence, and the column
// indices and A operands of the next DEPTH k-blocks are always in flight, also
// across row boundaries.  B traffic from L2 is K * TB bytes per workgroup
// instead of N bytes per nonzero vector, so the kernel pays off when a
// workgroup's rows together touch most rows of B.
#pragma once

#include "spmm_kernel.hpp"

namespace mc {

constexpr int kStatMaxRowsPerWave = 4;
// LDS budget of the B slice (K + 1 rows of TB bytes), below gfx950's 160 KiB.
constexpr int kStatLdsBytes = 160 * 1024 - 1024;

template <int V, int AB, bool SGN, int TB, int WPB, int DEPTH>
__global__ __launch_bounds__(64 * WPB) void spmm_stationary_kernel(SpmmArgs args, int row_groups) {
    using T = Pair<AB, 8, SGN>;
    constexpr int P = T::P;
    constexpr int ROWS = P * V;
    constexpr int RT = (ROWS + 15) / 16;
    constexpr int NCG = TB / 16;                 // 16-column MFMA groups per tile
    constexpr int OB = T::OB;
    constexpr int NT = 64 * WPB;
    constexpr u32 OOB = 0x80000000u;
    constexpr int HOME = V < 4 ? V : 4;
    constexpr int MAXR = kStatMaxRowsPerWave;
    extern __shared__ __attribute__((aligned(16))) unsigned char bslice[];

    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int r = lane & 15, g = lane >> 4;
    const int rg = blockIdx.x % row_groups;
    const int tile = blockIdx.x / row_groups;
    const int K = args.k;
    const int N = args.n;

    // ---- this wave's rows (of the workgroup's rows rg, rg + row_groups, ...).
    // Two cursors walk them in the same order: the issue cursor DEPTH k-blocks
    // ahead of the compute cursor.  Rows sit in shift queues so that every
    // array access has a constant index (no private-memory arrays).
    int qm[MAXR], qs[MAXR], qe[MAXR];
    bool qlive[MAXR];
    int total = 0;
#pragma unroll
    for (int u = 0; u < MAXR; ++u) {
        const int q = wave + WPB * u;                      // row of the workgroup
        const int slot = rg + row_groups * q;               // row of the launch order
        const bool live = slot < args.m_vec;
        const int m = live ? (args.row_indices ? args.row_indices[slot] : slot) : 0;
        const int s = live ? args.row_offsets[2 * m] : 0;
        const int e = live ? args.row_offsets[2 * m + 1] : 0;
        qlive[u] = live;
        qm[u] = m;
        qs[u] = s;
        qe[u] = s + ((e - s + 15) & ~15);                  // aligned end
        total += (qe[u] - qs[u] + 63) >> 6;
    }

    struct Cursor {
        int q_m[MAXR], q_s[MAXR], q_e[MAXR];
        int m = 0, s = 0, e = 0, kb = 0, n = 0;
        // moves to the next row that has k-blocks left (no-op while the current has)
        __device__ __forceinline__ void advance() {
#pragma unroll
            for (int u = 0; u < MAXR; ++u) {
                if (kb < n) break;
                m = q_m[0]; s = q_s[0]; e = q_e[0];
                n = (e - s + 63) >> 6;
                kb = 0;
#pragma unroll
                for (int w = 0; w + 1 < MAXR; ++w) { q_m[w] = q_m[w + 1]; q_s[w] = q_s[w + 1]; q_e[w] = q_e[w + 1]; }
                q_m[MAXR - 1] = 0; q_s[MAXR - 1] = 0; q_e[MAXR - 1] = 0;
            }
        }
    };
    Cursor ic, cc;
#pragma unroll
    for (int u = 0; u < MAXR; ++u) {
        ic.q_m[u] = cc.q_m[u] = qm[u];
        ic.q_s[u] = cc.q_s[u] = qs[u];
        ic.q_e[u] = cc.q_e[u] = qe[u];
    }

    const __amdgpu_buffer_rsrc_t ra = make_rsrc(args.values, 0x7FFFFFFFu);
    const __amdgpu_buffer_rsrc_t rc = make_rsrc(args.col_indices, 0x7FFFFFFFu);
    const __amdgpu_buffer_rsrc_t rb = make_rsrc(args.rhs, args.rhs_bytes);

    int a_p[RT], a_v[RT];
    bool a_live[RT];
#pragma unroll
    for (int t = 0; t < RT; ++t) {
        const int R = 16 * t + r;
        a_live[t] = R < ROWS;
        a_p[t] = R / V;
        a_v[t] = R % V;
    }

    struct Stage {
        int idx0, idx1;
        v4i a[RT];
    };
    // column indices of slots s0 + r/2 and s0 + 8 + r/2, and the A operand,
    // of the issue cursor's next k-block
    auto issue = [&](Stage &st) {
        ic.advance();
        const bool ok = ic.kb < ic.n;
        const int s0 = ic.s + 64 * ic.kb + 16 * g;
        const bool lok = ok && s0 < ic.e;
        if (ok) ++ic.kb;
        const u32 o0 = lok ? (u32)(s0 + (r >> 1)) * 4u : OOB;
        const int i0 = (int)buf_load_b32(rc, o0);
        const int i1 = (int)buf_load_b32(rc, lok ? o0 + 32u : OOB);
        st.idx0 = lok ? i0 : -1;
        st.idx1 = lok ? i1 : -1;
#pragma unroll
        for (int t = 0; t < RT; ++t) st.a[t] = load_a_operand<V, AB, 8, SGN>(ra, s0, a_p[t], a_v[t], lok && a_live[t]);
    };

    Stage ring[DEPTH];
#pragma unroll
    for (int d = 0; d < DEPTH; ++d) {
        issue(ring[d]);
        __builtin_amdgcn_sched_barrier(0);      // the loop expects this issue order
    }

    // ---- the B slice: rows 0..K-1 of columns [tile*TB, tile*TB + TB), then a zero row
    {
        constexpr int CPR = TB / 16;
        const int chunks = (K + 1) * CPR;
        const u32 col0 = (u32)tile * TB;
        for (int c0 = threadIdx.x; c0 < chunks; c0 += 4 * NT) {
            v4i x[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int c = c0 + u * NT;
                const int row = c / CPR, part = c % CPR;
                const u32 off = (c < chunks && row < K) ? (u32)row * (u32)N + col0 + 16u * part : OOB;
                x[u] = buf_load_b128(rb, off);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int c = c0 + u * NT;
                if (c < chunks) *reinterpret_cast<v4i *>(bslice + (size_t)c * 16) = x[u];
            }
        }
    }
    __syncthreads();

    v4i acc[RT][NCG];
    int rsum[RT];
    auto reset = [&]() {
#pragma unroll
        for (int t = 0; t < RT; ++t) {
            rsum[t] = 0;
#pragma unroll
            for (int cg = 0; cg < NCG; ++cg) acc[t][cg] = v4i{0, 0, 0, 0};
        }
    };
    reset();

    const u32 seg = 8u * (u32)(r & 1);
    auto lds_row = [&](int idx) { return (u32)((unsigned)idx < (unsigned)K ? idx : K) * (u32)TB; };

    // plane recombination and store of one finished row (all lanes take part)
    auto finish_row = [&](int m) {
        int rs_home[HOME];
#pragma unroll
        for (int i = 0; i < HOME; ++i) rs_home[i] = 0;
        if constexpr (OB != 0) {
            int s_full[RT];
#pragma unroll
            for (int t = 0; t < RT; ++t) {
                int s = rsum[t];
                s += lane_read(s, lane ^ 16);
                s += lane_read(s, lane ^ 32);
                s_full[t] = s;
            }
#pragma unroll
            for (int i = 0; i < HOME; ++i) {
                const int v = 4 * g + i;
#pragma unroll
                for (int p = 0; p < P; ++p) {
                    const int R = p * V + v;
                    int sv = 0;
#pragma unroll
                    for (int t = 0; t < RT; ++t) {
                        const int got = lane_read(s_full[t], R & 15);
                        if ((R >> 4) == t) sv = got;
                    }
                    rs_home[i] += T::weight(p) * sv;
                }
            }
        }
        int res[NCG][HOME];
#pragma unroll
        for (int cg = 0; cg < NCG; ++cg)
#pragma unroll
            for (int i = 0; i < HOME; ++i) {
                int vsum = OB * rs_home[i];
#pragma unroll
                for (int p = 0; p < P; ++p) {
                    int t, dg, di;
                    if constexpr (V == 8) { t = p / 2; dg = 2 * (p % 2); di = 0; }
                    else if constexpr (V == 4) { t = p / 4; dg = p % 4; di = 0; }
                    else { t = p / 8; dg = (p % 8) / 2; di = 2 * (p % 2); }
                    int x = acc[t][cg][i + di];
                    if (dg != 0) x = lane_down(x, 16 * dg);
                    vsum += T::weight(p) * x;
                }
                res[cg][i] = vsum;
            }
        if (g < (V + 3) / 4) {
#pragma unroll
            for (int i = 0; i < HOME; ++i) {
                int *orow = args.out + (size_t)(m * V + 4 * g + i) * (size_t)N + (size_t)tile * TB + r;
#pragma unroll
                for (int cg = 0; cg < NCG; ++cg) __builtin_nontemporal_store(res[cg][i], orow + 16 * cg);
            }
        }
    };

    auto compute = [&](const Stage &st) {
        cc.advance();
        const u32 a0 = lds_row(st.idx0) + seg, a1 = lds_row(st.idx1) + seg;
#pragma unroll
        for (int cg = 0; cg < NCG; ++cg) {
            const v2i lo = lds_read_tr8(bslice, a0 + 16u * cg);
            const v2i hi = lds_read_tr8(bslice, a1 + 16u * cg);
            v4i bop = v4i{lo[0], lo[1], hi[0], hi[1]};
            if constexpr (!SGN) bop ^= v4i{(int)0x80808080u, (int)0x80808080u, (int)0x80808080u, (int)0x80808080u};
#pragma unroll
            for (int t = 0; t < RT; ++t) acc[t][cg] = mfma_i8(st.a[t], bop, acc[t][cg]);
        }
        if constexpr (OB != 0) {
#pragma unroll
            for (int t = 0; t < RT; ++t) rsum[t] += sum_bytes(st.a[t]);
        }
        if (++cc.kb == cc.n) {               // last k-block of the row
            finish_row(cc.m);
            reset();
        }
    };

    // rows without nonzero vectors never reach finish_row: their outputs are zero
    if (g < (V + 3) / 4) {
#pragma unroll
        for (int u = 0; u < MAXR; ++u) {
            if (qlive[u] && qs[u] == qe[u]) {
#pragma unroll
                for (int i = 0; i < HOME; ++i) {
                    int *orow = args.out + (size_t)(qm[u] * V + 4 * g + i) * (size_t)N + (size_t)tile * TB + r;
#pragma unroll
                    for (int cg = 0; cg < NCG; ++cg) __builtin_nontemporal_store(0, orow + 16 * cg);
                }
            }
        }
    }

    // branch-free main loop while every stage holds a k-block, then the tail
    int base = 0;
    for (; base + DEPTH <= total; base += DEPTH) {
#pragma unroll
        for (int d = 0; d < DEPTH; ++d) {
            compute(ring[d]);
            issue(ring[d]);
            __builtin_amdgcn_sched_barrier(0);
        }
    }
#pragma unroll
    for (int d = 0; d < DEPTH; ++d)
        if (base + d < total) compute(ring[d]);
}

// Launch geometry of the stationary kernel for a shape (0 when the shape
// does not fit: N not a multiple of the tile, or the K rows not in LDS).
struct StatGeometry {
    int row_groups = 0;
    int tiles = 0;
    size_t lds = 0;
};

template <int TB, int WPB>
StatGeometry stationary_geometry(int m_vec, int n, int k, int rows_per_block) {
    StatGeometry geo;
    if (n % TB != 0 || m_vec <= 0) return geo;
    const size_t lds = (size_t)(k + 1) * TB;
    if (lds > (size_t)kStatLdsBytes) return geo;
    const int max_rows = kStatMaxRowsPerWave * WPB;
    const int rows = rows_per_block < 1 ? 1 : (rows_per_block > max_rows ? max_rows : rows_per_block);
    geo.row_groups = (m_vec + rows - 1) / rows;
    geo.tiles = n / TB;
    geo.lds = lds;
    return geo;
}

template <int V, int AB, bool SGN, int TB, int WPB, int D>
mc_status spmm_stationary_launch(SpmmArgs a, int rows_per_block, hipStream_t stream) {
    const StatGeometry geo = stationary_geometry<TB, WPB>(a.m_vec, a.n, a.k, rows_per_block);
    if (geo.row_groups == 0) return MC_EINVAL_SHAPE;
    const long long blocks = (long long)geo.row_groups * geo.tiles;
    if (blocks > 0x7fffffffLL) return MC_EINVAL_SHAPE;
    auto kern = spmm_stationary_kernel<V, AB, SGN, TB, WPB, D>;
    static bool attr_set = false;     // raising the dynamic-LDS limit is idempotent
    if (!attr_set) {
        if (hipFuncSetAttribute(reinterpret_cast<const void *>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                kStatLdsBytes) != hipSuccess)
            return check_launch();
        attr_set = true;
    }
    hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(64 * WPB), geo.lds, stream, a, geo.row_groups);
    return check_launch();
}

}  // namespace mc
