// Quantized vector-sparse SpMM for gfx950: C[M x N] (int32) = A * B.
//
// A is Magicube's vector-CSR (V x 1 column vectors, rows padded to MMA_K
// slots, (start, end) pairs), with values in the reference's prepared layout
// (SpMM/SpMM/spmm_benchmark.cpp:238-377).  B is K x N row-major, packed
// 4 x int8 or 8 x int4 per int32 with lane 0 in the low bits.
//
// Work unit: one wave = one vector row x TILE_N columns x (1/SPLITK) of the
// row's nonzero vectors.  Per 64-slot k-block every lane
//   * loads the 16 column indices of its lane group (slots 16g..16g+15),
//   * gathers LW dwords of each of those 16 B rows with raw buffer loads
//     (padding slots, index -1, and bad indices fall outside the buffer and
//     read as zero without touching memory),
//   * byte-transposes 4x4 blocks with v_perm so every lane holds 16
//     reduction slots of one column, and issues v_mfma_i32_16x16x64_i8.
//
// Operand digits.  The MFMA is signed int8 x int8, so every operand is a sum
// of int8 digit planes with power-of-two weights; A planes stack as MFMA rows
// (row = plane*V + v, padded to 16-row tiles), and the planes are recombined
// by shifted int32 adds in the epilogue.  B digits may carry a constant offset
// OB (u8 -> b^0x80 = b-128, signed i4 -> nibble^8 = b+8), corrected with
// OB * (row sum of A), computed on the A operand with v_dot4.  16-bit B
// columns are two byte digits (high x256, low x1), each its own MFMA column
// group, recombined in the epilogue.  12-bit A sits in 16-bit containers.
//
//   pair    signed                          unsigned
//   L8-R8   A s8                             A = 16*hi4 + lo4, B = d + 128
//   L16-R8  A = 256*s8 + 16*u4 + u4          A = 4 nibbles,  B = d + 128
//   L12-R8  A = 16*s8 + u4                   A = 3 nibbles,  B = d + 128
//   L16-R16 A as L16-R8, B = 256*s8 + d + 128  A = 4 nibbles, B = 256*d' + d + 32896
//   L8-R4   A s8 (from the two nibble        A = 16*hi4 + lo4
//           planes), B = d - 8
//   L16-R4  A = 256*s8 + 16*u4 + u4, B = d - 8   A = 4 nibbles, B u4
//   L12-R4  A = 16*s8 + u4, B = d - 8        A = 3 nibbles,  B u4
//   L4-R4   A s4->s8, B = d - 8              A u4, B u4
//
// The output is the exact sum modulo 2^32.
#pragma once

#include "mc_device.hpp"

namespace mc {

struct SpmmArgs {
    int m_vec;
    int n;
    int k;
    const int *row_indices;     // row swizzle (nullptr: identity)
    const int *row_offsets;     // 2*m_vec (start, end) pairs
    const int *col_indices;     // padded with -1; shuffled in 8-groups for 4-bit B
    const unsigned char *values;
    const unsigned char *rhs;
    int *out;
    int tiles_n;                // ceil(n / TILE_N)
    unsigned rhs_bytes;         // k * n * BB / 8  (< 2^31)
    // batched launch (blockIdx.y = matrix b of one shared mask, gather
    // kernel only): matrix b's values / rhs start values_stride /
    // rhs_stride bytes after matrix b-1's, its output out_stride int32
    // after; 0 for a single matrix
    long long values_stride;
    long long rhs_stride;
    long long out_stride;
    int batch;                  // matrices in the launch (gridDim.y), >= 1
};

template <int AB, int BB, bool SGN> struct Pair;

template <bool SGN> struct Pair<8, 8, SGN> {
    static constexpr int MMA_K = 16, P = SGN ? 1 : 2, OB = SGN ? 0 : 128;
    __device__ static constexpr int weight(int p) { return SGN ? 1 : (p == 0 ? 16 : 1); }
};
template <bool SGN> struct Pair<16, 8, SGN> {
    static constexpr int MMA_K = 16, P = SGN ? 3 : 4, OB = SGN ? 0 : 128;
    __device__ static constexpr int weight(int p) {
        return SGN ? (p == 0 ? 256 : p == 1 ? 16 : 1) : (p == 0 ? 4096 : p == 1 ? 256 : p == 2 ? 16 : 1);
    }
};
template <bool SGN> struct Pair<8, 4, SGN> {
    static constexpr int MMA_K = 32, P = SGN ? 1 : 2, OB = SGN ? -8 : 0;
    __device__ static constexpr int weight(int p) { return SGN ? 1 : (p == 0 ? 16 : 1); }
};
template <bool SGN> struct Pair<4, 4, SGN> {
    static constexpr int MMA_K = 32, P = 1, OB = SGN ? -8 : 0;
    __device__ static constexpr int weight(int) { return 1; }
};

// 12-bit A: unsigned as three nibbles; signed as the arithmetic top byte
// (bits 4..11, x16) and the low nibble.
template <bool SGN> struct Pair<12, 8, SGN> {
    static constexpr int MMA_K = 16, P = SGN ? 2 : 3, OB = SGN ? 0 : 128;
    __device__ static constexpr int weight(int p) { return SGN ? (p == 0 ? 16 : 1) : (p == 0 ? 256 : p == 1 ? 16 : 1); }
};
template <bool SGN> struct Pair<12, 4, SGN> {
    static constexpr int MMA_K = 32, P = SGN ? 2 : 3, OB = SGN ? -8 : 0;
    __device__ static constexpr int weight(int p) { return SGN ? (p == 0 ? 16 : 1) : (p == 0 ? 256 : p == 1 ? 16 : 1); }
};
template <bool SGN> struct Pair<16, 4, SGN> {
    static constexpr int MMA_K = 32, P = SGN ? 3 : 4, OB = SGN ? -8 : 0;
    __device__ static constexpr int weight(int p) {
        return SGN ? (p == 0 ? 256 : p == 1 ? 16 : 1) : (p == 0 ? 4096 : p == 1 ? 256 : p == 2 ? 16 : 1);
    }
};
// 16-bit B: the offset covers both byte digits (unsigned: 128*256 + 128;
// signed: the low byte only).
template <bool SGN> struct Pair<16, 16, SGN> {
    static constexpr int MMA_K = 16, P = SGN ? 3 : 4, OB = SGN ? 128 : 32896;
    __device__ static constexpr int weight(int p) {
        return SGN ? (p == 0 ? 256 : p == 1 ? 16 : 1) : (p == 0 ? 4096 : p == 1 ? 256 : p == 2 ? 16 : 1);
    }
};

// Where the A operand bytes come from: the values buffer itself, or a copy
// of one row chunk's values staged in LDS (byte offsets stay those of the
// buffer; an offset with bit 31 set reads as zeros in both).
struct GlobalA {
    __amdgpu_buffer_rsrc_t r;
    __device__ __forceinline__ v4i b128(u32 off) const { return buf_load_b128(r, off); }
    __device__ __forceinline__ v2i b64(u32 off) const { return buf_load_b64(r, off); }
};
struct LdsA {
    const unsigned char *lds;   // chunk copy, starting at buffer offset `base`
    u32 base;
    __device__ __forceinline__ v4i b128(u32 off) const {
        const bool in = off < 0x80000000u;
        const v4i x = *reinterpret_cast<const v4i *>(lds + (in ? off - base : 0u));
        return in ? x : v4i{0, 0, 0, 0};
    }
    __device__ __forceinline__ v2i b64(u32 off) const {
        const bool in = off < 0x80000000u;
        const v2i x = *reinterpret_cast<const v2i *>(lds + (in ? off - base : 0u));
        return in ? x : v2i{0, 0};
    }
};

// A operand of one lane: plane p, vector lane v, reduction slots
// [slot0, slot0 + 16).  Returns zeros when !ok (buffer offset out of range).
template <int V, int AB, int BB, bool SGN, class Src>
__device__ __forceinline__ v4i load_a_operand(const Src &va, int slot0, int p, int v, bool ok) {
    const u32 OOB = 0x80000000u;
    if constexpr (BB != 4) {
        // groups of 16 slots; [V][16] bytes per byte plane
        const int grp = slot0 >> 4;
        if constexpr (AB == 8) {
            const u32 off = ok ? (u32)(grp * 16 * V + 16 * v) : OOB;
            v4i x = va.b128(off);
            if constexpr (SGN) {
                return x;
            } else {
                const int sh = (p == 0) ? 4 : 0;
                v4i y;
#pragma unroll
                for (int q = 0; q < 4; ++q) y[q] = (int)(((u32)x[q] >> sh) & 0x0F0F0F0Fu);
                return y;
            }
        } else if constexpr (AB == 16) {  // low-byte plane then high-byte plane per group
            const bool hi = SGN ? (p == 0) : (p < 2);
            const u32 off = ok ? (u32)(grp * 32 * V + (hi ? 16 * V : 0) + 16 * v) : OOB;
            v4i x = va.b128(off);
            if (SGN && p == 0) return x;
            const int sh = SGN ? (p == 1 ? 4 : 0) : ((p & 1) == 0 ? 4 : 0);
            v4i y;
#pragma unroll
            for (int q = 0; q < 4; ++q) y[q] = (int)(((u32)x[q] >> sh) & 0x0F0F0F0Fu);
            return y;
        } else {  // AB == 12: the 16-bit layout; the high-byte plane holds bits 8..11
            const u32 base = (u32)(grp * 32 * V + 16 * v);
            if constexpr (SGN) {
                // digit 0 = bits 4..11 (signed byte), digit 1 = bits 0..3
                v4i lo = va.b128(ok ? base : OOB);
                v4i hi = va.b128((ok && p == 0) ? base + 16 * V : OOB);
                v4i y;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const u32 top = (((u32)hi[q] << 4) & 0xF0F0F0F0u) | (((u32)lo[q] >> 4) & 0x0F0F0F0Fu);
                    y[q] = (int)(p == 0 ? top : ((u32)lo[q] & 0x0F0F0F0Fu));
                }
                return y;
            } else {
                // digits: bits 8..11 (high plane), 4..7, 0..3 (low plane)
                v4i x = va.b128(ok ? base + (p == 0 ? 16 * V : 0) : OOB);
                const int sh = (p == 1) ? 4 : 0;
                v4i y;
#pragma unroll
                for (int q = 0; q < 4; ++q) y[q] = (int)(((u32)x[q] >> sh) & 0x0F0F0F0Fu);
                return y;
            }
        }
    } else {
        // groups of 32 slots; the lane's 16 slots are half h of the group.
        // Slot order follows the 8-group column shuffle (evens then odds),
        // which is exactly [even nibbles(x0), odd(x0), even(x1), odd(x1)].
        const int grp = slot0 >> 5, h = (slot0 >> 4) & 1;
        if constexpr (AB == 4) {
            const u32 off = ok ? (u32)(grp * 16 * V + 16 * v + 8 * h) : OOB;
            v2i x = va.b64(off);
            u32 e0 = nib_even((u32)x[0]), o0 = nib_odd((u32)x[0]);
            u32 e1 = nib_even((u32)x[1]), o1 = nib_odd((u32)x[1]);
            if constexpr (SGN) { e0 = sext4_bytes(e0); o0 = sext4_bytes(o0); e1 = sext4_bytes(e1); o1 = sext4_bytes(o1); }
            return v4i{(int)e0, (int)o0, (int)e1, (int)o1};
        } else if constexpr (AB == 8) {  // low-nibble plane then high-nibble plane
            const u32 base = (u32)(grp * 32 * V + 16 * v + 8 * h);
            if constexpr (SGN) {
                v2i lo = va.b64(ok ? base : OOB);
                v2i hi = va.b64(ok ? base + 16 * V : OOB);
                u32 e0 = nib_even((u32)lo[0]) | (nib_even((u32)hi[0]) << 4);
                u32 o0 = nib_odd((u32)lo[0]) | (nib_odd((u32)hi[0]) << 4);
                u32 e1 = nib_even((u32)lo[1]) | (nib_even((u32)hi[1]) << 4);
                u32 o1 = nib_odd((u32)lo[1]) | (nib_odd((u32)hi[1]) << 4);
                return v4i{(int)e0, (int)o0, (int)e1, (int)o1};
            } else {
                v2i x = va.b64(ok ? base + (p == 0 ? 16 * V : 0) : OOB);
                return v4i{(int)nib_even((u32)x[0]), (int)nib_odd((u32)x[0]),
                           (int)nib_even((u32)x[1]), (int)nib_odd((u32)x[1])};
            }
        } else {
            // AB == 12 or 16: nibble planes q = 0..AB/4-1 (bits 4q..4q+3) at
            // 16*V*q in a 64*V-byte group (the 12-bit layout leaves plane 3 empty)
            constexpr int TOP = AB / 4 - 1;
            const u32 base = (u32)(grp * 64 * V + 16 * v + 8 * h);
            if constexpr (SGN) {
                // digit 0 = planes TOP (high nibble) and TOP-1 as one signed
                // byte; digit p >= 1 = plane TOP-1-p
                const int plo = p == 0 ? TOP - 1 : TOP - 1 - p;
                v2i lo = va.b64(ok ? base + (u32)(16 * V * plo) : OOB);
                v2i hi = va.b64((ok && p == 0) ? base + (u32)(16 * V * TOP) : OOB);
                u32 e0 = nib_even((u32)lo[0]) | (nib_even((u32)hi[0]) << 4);
                u32 o0 = nib_odd((u32)lo[0]) | (nib_odd((u32)hi[0]) << 4);
                u32 e1 = nib_even((u32)lo[1]) | (nib_even((u32)hi[1]) << 4);
                u32 o1 = nib_odd((u32)lo[1]) | (nib_odd((u32)hi[1]) << 4);
                return v4i{(int)e0, (int)o0, (int)e1, (int)o1};
            } else {
                v2i x = va.b64(ok ? base + (u32)(16 * V * (TOP - p)) : OOB);
                return v4i{(int)nib_even((u32)x[0]), (int)nib_odd((u32)x[0]),
                           (int)nib_even((u32)x[1]), (int)nib_odd((u32)x[1])};
            }
        }
    }
}
template <int V, int AB, int BB, bool SGN>
__device__ __forceinline__ v4i load_a_operand(__amdgpu_buffer_rsrc_t va, int slot0, int p, int v, bool ok) {
    return load_a_operand<V, AB, BB, SGN>(GlobalA{va}, slot0, p, v, ok);
}

template <int LW>
__device__ __forceinline__ void load_b_row(__amdgpu_buffer_rsrc_t rb, u32 off, u32 (&dst)[LW]) {
    if constexpr (LW == 1) {
        dst[0] = buf_load_b32(rb, off);
    } else if constexpr (LW == 2) {
        v2i x = buf_load_b64(rb, off);
        dst[0] = (u32)x[0]; dst[1] = (u32)x[1];
    } else {
        v4i x = buf_load_b128(rb, off);
        dst[0] = (u32)x[0]; dst[1] = (u32)x[1]; dst[2] = (u32)x[2]; dst[3] = (u32)x[3];
    }
}

// Gather offsets of one row chunk, staged in LDS per workgroup.
constexpr int kStageSlots = 1024;
// B is limited to 1 GiB: offsets of real rows are < 2^30, and adding this
// constant (to a staged offset or to a padding marker of the same value)
// always lands in [2^30, 2^31), outside the buffer.
constexpr u32 kOobAdd = 0x40000000u;

// WPS: waves per SIMD the register allocation must allow (1: no constraint).
template <int V, int AB, int BB, bool SGN, int LW, int TPB, int SPLITK, int DEPTH, int WPS>
__global__ __launch_bounds__(64 * TPB * SPLITK) __attribute__((amdgpu_waves_per_eu(WPS))) void spmm_kernel(
    SpmmArgs args) {
    using T = Pair<AB, BB, SGN>;
    constexpr int P = T::P;
    constexpr int ROWS = P * V;
    constexpr int RT = (ROWS + 15) / 16;
    constexpr int CPD = 32 / BB;            // columns per B dword
    // 16-column MFMA groups per B dword: one per column, two per 16-bit
    // column (low digit at even, high digit at odd index)
    constexpr int NV = BB == 4 ? 8 : 4;
    constexpr int NMMA = LW * NV;           // accumulator groups per wave
    constexpr int NCOL = LW * CPD;          // output columns per lane
    constexpr int TILE_N = 16 * NCOL;
    constexpr int MMA_K = T::MMA_K;
    constexpr int OB = T::OB;
    constexpr int NT = 64 * TPB * SPLITK;
    constexpr u32 OOB = 0x80000000u;
    constexpr int HOME = V < 4 ? V : 4;      // output rows held per home lane
    __shared__ __attribute__((aligned(16))) u32 s_off[kStageSlots];
    // With several waves per row, the chunk's A values are staged in LDS once
    // per workgroup instead of being loaded by every wave: a G-slot group
    // holds G * V * container-bits / 8 bytes, and rows start on a group.
    // Measured on MI355X: 1-5 % faster for 8- and 16-bit B, 2 % slower for
    // 4-bit B (whose A loads are 8 bytes per lane), which keeps global loads.
    constexpr bool kShareA = TPB > 1 && BB != 4;
    constexpr int AG = BB == 4 ? 32 : 16;
    constexpr int AGB = AG * V * (AB == 12 ? 16 : AB) / 8;
    __shared__ __attribute__((aligned(16))) unsigned char s_a[kShareA ? kStageSlots / AG * AGB : 16];

    const int lane = threadIdx.x & 63;
    // wave-uniform in an SGPR, so every loop bound below is scalar
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int r = lane & 15, g = lane >> 4;
    const int tile_local = wave % TPB;
    const int split = wave / TPB;
    const int tile_groups = (args.tiles_n + TPB - 1) / TPB;
    const int row_slot = blockIdx.x / tile_groups;
    const int tile = (blockIdx.x % tile_groups) * TPB + tile_local;

    const int m = args.row_indices ? args.row_indices[row_slot] : row_slot;
    const int start = args.row_offsets[2 * m];
    const int end = args.row_offsets[2 * m + 1];
    const int aend = start + ((end - start + MMA_K - 1) / MMA_K) * MMA_K;

    const int n_dw = args.n / CPD;                        // dwords per B row
    const int col_dw = tile * (TILE_N / CPD) + r * LW;    // first dword of this lane
    const bool col_ok = (tile < args.tiles_n) && (col_dw < n_dw);
    const u32 row_stride = (u32)n_dw * 4u;
    // lanes past the last column add kOobAdd: real rows then read out of range
    const u32 coff = col_ok ? (u32)col_dw * 4u : kOobAdd;

    const long long bz = blockIdx.y;
    const __amdgpu_buffer_rsrc_t rb = make_rsrc(args.rhs + bz * args.rhs_stride, args.rhs_bytes);
    const __amdgpu_buffer_rsrc_t ra = make_rsrc(args.values + bz * args.values_stride, 0x7FFFFFFFu);
    int *const out = args.out + bz * args.out_stride;

    // plane / vector lane of this lane's A row in each row tile
    int a_p[RT], a_v[RT];
    bool a_live[RT];
#pragma unroll
    for (int t = 0; t < RT; ++t) {
        const int R = 16 * t + r;
        a_live[t] = R < ROWS;
        a_p[t] = R / V;
        a_v[t] = R % V;
    }

    v4i acc[RT][NMMA];
#pragma unroll
    for (int t = 0; t < RT; ++t)
#pragma unroll
        for (int j = 0; j < NMMA; ++j) acc[t][j] = v4i{0, 0, 0, 0};
    int rsum[RT];
#pragma unroll
    for (int t = 0; t < RT; ++t) rsum[t] = 0;

    auto compute = [&](u32 (&braw)[16][LW], const v4i (&aop)[RT]) {
#pragma unroll
        for (int d = 0; d < LW; ++d) {
            v4i bop[NV];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                u32 o[4];
                if constexpr (BB != 4) {
                    // unsigned bytes become signed digits b - 128 (OB = +128);
                    // 16-bit B: bytes 0/2 of a dword are the low digits of its
                    // two columns, 1/3 the high digits, so o[b] is column b/2's
                    // digit b%2
                    const u32 x = BB == 8 ? (SGN ? 0u : 0x80808080u) : (SGN ? 0x00800080u : 0x80808080u);
                    transpose4x4(braw[4 * q][d] ^ x, braw[4 * q + 1][d] ^ x, braw[4 * q + 2][d] ^ x,
                                 braw[4 * q + 3][d] ^ x, o);
#pragma unroll
                    for (int b = 0; b < 4; ++b) bop[b][q] = (int)o[b];
                } else {
                    // signed nibbles become digits b + 8 (OB = -8)
                    const u32 x = SGN ? 0x88888888u : 0u;
                    transpose4x4(braw[4 * q][d] ^ x, braw[4 * q + 1][d] ^ x, braw[4 * q + 2][d] ^ x,
                                 braw[4 * q + 3][d] ^ x, o);
#pragma unroll
                    for (int b = 0; b < 4; ++b) {
                        bop[2 * b][q] = (int)nib_even(o[b]);
                        bop[2 * b + 1][q] = (int)nib_odd(o[b]);
                    }
                }
            }
#pragma unroll
            for (int t = 0; t < RT; ++t)
#pragma unroll
                for (int c = 0; c < NV; ++c) acc[t][d * NV + c] = mfma_i8(aop[t], bop[c], acc[t][d * NV + c]);
        }
        if constexpr (OB != 0) {
#pragma unroll
            for (int t = 0; t < RT; ++t) rsum[t] += sum_bytes(aop[t]);
        }
    };

    // The row's slots are processed in chunks of kStageSlots.  Per chunk the
    // workgroup stages each slot's B byte offset (index * row stride, OOB for
    // padding) in LDS once; each wave then keeps DEPTH k-blocks of gathers in
    // flight, reading its lane group's 16 offsets with LDS broadcasts.
    for (int c0 = start; c0 < aend; c0 += kStageSlots) {
        const int cn = min(aend - c0, kStageSlots);
        const int cblk = (cn + 63) >> 6;
        __syncthreads();
        for (int i = threadIdx.x; i < cblk * 64; i += NT) {
            u32 o = kOobAdd;
            if (i < cn) {
                const int idx = args.col_indices[c0 + i];
                if (idx >= 0) o = (u32)idx * row_stride;
            }
            s_off[i] = o;
        }
        // the chunk's cn slots (a multiple of the group size) of A values
        const u32 a_base = (u32)(c0 / AG) * (u32)AGB;
        if constexpr (kShareA) {
            const int n16 = cn / AG * AGB / 16;
            for (int i = threadIdx.x; i < n16; i += NT)
                *reinterpret_cast<v4i *>(s_a + 16 * i) = buf_load_b128(ra, a_base + 16u * (u32)i);
        }
        __syncthreads();

        // Branch-free issue: the LDS block index is clamped into the chunk and
        // validity rides on one additive offset (B is < 1 GiB, so adding
        // kOobAdd to any staged offset lands outside the buffer, no wrap).
        auto issue = [&](int kb, u32 (&br)[16][LW], v4i (&ar)[RT]) {
            const bool ok = kb < cblk;
            const int lb = ok ? kb : cblk - 1;
            const u32 kadd = ok ? coff : kOobAdd;
            const int slot0 = c0 + lb * 64 + 16 * g;
            const bool aok = ok && (slot0 < aend);
            const v4i *so = reinterpret_cast<const v4i *>(s_off + lb * 64 + 16 * g);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const v4i o4 = so[q];
#pragma unroll
                for (int e = 0; e < 4; ++e) load_b_row<LW>(rb, (u32)o4[e] + kadd, br[4 * q + e]);
            }
#pragma unroll
            for (int t = 0; t < RT; ++t) {
                if constexpr (kShareA)
                    ar[t] = load_a_operand<V, AB, BB, SGN>(LdsA{s_a, a_base}, slot0, a_p[t], a_v[t], aok && a_live[t]);
                else
                    ar[t] = load_a_operand<V, AB, BB, SGN>(ra, slot0, a_p[t], a_v[t], aok && a_live[t]);
            }
        };

        u32 ring[DEPTH][16][LW];
        v4i aring[DEPTH][RT];
#pragma unroll
        for (int d = 0; d < DEPTH; ++d) {
            issue(split + d * SPLITK, ring[d], aring[d]);
            __builtin_amdgcn_sched_barrier(0);      // the loop expects this issue order
        }
        // Main loop: every stage of the ring holds a real k-block, no branches
        // inside, so the wait counts keep DEPTH-1 stages of gathers in flight
        // across the back edge.  Then the last partial round.
        int base = split;
        for (; base + (DEPTH - 1) * SPLITK < cblk; base += DEPTH * SPLITK) {
#pragma unroll
            for (int d = 0; d < DEPTH; ++d) {
                compute(ring[d], aring[d]);
                issue(base + d * SPLITK + DEPTH * SPLITK, ring[d], aring[d]);
                // keep stage d's work here: hoisting the next stages' transposes
                // above this point would wait for their gathers too
                __builtin_amdgcn_sched_barrier(0);
            }
        }
#pragma unroll
        for (int d = 0; d < DEPTH; ++d)
            if (base + d * SPLITK < cblk) compute(ring[d], aring[d]);
    }

    // ---- epilogue: recombine planes (shifted adds), offset correction ----
    // Output row v lives in home lane group v/4 (element v%4).  Plane p of
    // row v sits in MFMA row R = p*V + v: tile R/16, lane group (R%16)/4,
    // element R%4.
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
            s_full[t] = s;  // row 16t + r
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

    // sums mod 2^32 in unsigned arithmetic
    u32 comb[NMMA][HOME];
#pragma unroll
    for (int j = 0; j < NMMA; ++j)
#pragma unroll
        for (int i = 0; i < HOME; ++i) {
            u32 vsum = 0;
#pragma unroll
            for (int p = 0; p < P; ++p) {
                int t, dg, di;
                if constexpr (V == 8) { t = p / 2; dg = 2 * (p % 2); di = 0; }
                else if constexpr (V == 4) { t = p / 4; dg = p % 4; di = 0; }
                else { t = p / 8; dg = (p % 8) / 2; di = 2 * (p % 2); }
                int x = acc[t][j][i + di];
                if (dg != 0) x = lane_down(x, 16 * dg);
                vsum += (u32)T::weight(p) * (u32)x;
            }
            comb[j][i] = vsum;
        }
    int res[NCOL][HOME];
#pragma unroll
    for (int c = 0; c < NCOL; ++c)
#pragma unroll
        for (int i = 0; i < HOME; ++i) {
            u32 x;
            if constexpr (BB == 16) x = (comb[2 * c + 1][i] << 8) + comb[2 * c][i];
            else x = comb[c][i];
            res[c][i] = (int)(x + (u32)OB * (u32)rs_home[i]);
        }

    const bool home = g < (V + 3) / 4;
    if constexpr (SPLITK > 1) {
        // waves with split > 0 hand their partial tile to split 0 via LDS
        __shared__ int part[(SPLITK > 1 ? SPLITK - 1 : 1) * TPB][V][TILE_N];
        if (split > 0 && home) {
#pragma unroll
            for (int j = 0; j < NCOL; ++j) {
                const int col = r * NCOL + j;
#pragma unroll
                for (int i = 0; i < HOME; ++i) part[(split - 1) * TPB + tile_local][4 * g + i][col] = res[j][i];
            }
        }
        __syncthreads();
        if (split > 0) return;
        if (home) {
#pragma unroll
            for (int s = 1; s < SPLITK; ++s)
#pragma unroll
                for (int j = 0; j < NCOL; ++j) {
                    const int col = r * NCOL + j;
#pragma unroll
                    for (int i = 0; i < HOME; ++i)
                        res[j][i] = (int)((u32)res[j][i] + (u32)part[(s - 1) * TPB + tile_local][4 * g + i][col]);
                }
        }
    }

    if (!home || !col_ok) return;
    // lane r owns dwords [col_dw, col_dw + LW): NCOL consecutive columns
#pragma unroll
    for (int i = 0; i < HOME; ++i) {
        const int v = 4 * g + i;
        int *orow = out + (size_t)(m * V + v) * (size_t)args.n + (size_t)col_dw * CPD;
        if constexpr (NCOL % 4 == 0) {
#pragma unroll
            for (int c4 = 0; c4 < NCOL; c4 += 4) {
                v4i o = v4i{res[c4][i], res[c4 + 1][i], res[c4 + 2][i], res[c4 + 3][i]};
                __builtin_nontemporal_store(o, reinterpret_cast<v4i *>(orow + c4));
            }
        } else {
#pragma unroll
            for (int c2 = 0; c2 < NCOL; c2 += 2) {
                v2i o = v2i{res[c2][i], res[c2 + 1][i]};
                __builtin_nontemporal_store(o, reinterpret_cast<v2i *>(orow + c2));
            }
        }
    }
}

}  // namespace mc
