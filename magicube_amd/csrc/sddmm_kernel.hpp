// Quantized vector-sparse SDDMM for gfx950:
//   out[j][v] = sum_k lhs[i*V + v][k] * rhs[col_j][k]   for every nonzero slot j
// of vector row i (Magicube's wmmaSddmm, SDDMM/SDDMM/src/wmma_sddmm.cu:15-198).
//
// lhs is M x K and rhs is N x K, row-major, packed 2 x int16, 4 x int8 or
// 8 x int4 per int32 (lane 0 low).  Both operands are contiguous along K, so MFMA operands
// are plain 16-byte (8-bit) or 8-byte (4-bit) loads: lane (r, g) takes the
// 16 k-values [64*kc + 16g, +16) of lhs row r (A operand) and of the gathered
// rhs row of nonzero r (B operand) — no transposes.  One MFMA covers 16
// nonzeros x V rows x 64 k.
//
// Digits as in the SpMM kernel: unsigned 8-bit lhs splits into two nibble
// planes (MFMA rows plane*V + v); unsigned 8-bit rhs uses d = b ^ 0x80 with
// offset +128, signed 4-bit rhs d = b ^ 8 with offset -8; the offset is
// corrected with the lhs row sum over K.  16-bit operands are two byte
// digits on both sides (high x256, low x1; unsigned bytes and signed low
// bytes as b ^ 0x80), so lhs = lhs' + OA and rhs = rhs' + OB, and
//   sum lhs*rhs = sum lhs'*rhs' + OB*sum lhs' + OA*sum rhs' + K*OA*OB,
// with sum rhs' taken per nonzero from the gathered operand.  The lhs digit
// planes stack as MFMA rows (high plane rows 0..V-1); the two rhs digits are
// two accumulators.  Results are exact modulo 2^32.
//
// K is any multiple of 64 (the reference's Tile_K loop,
// SDDMM/SDDMM/src/wmma_sddmm.cu:77,174,261): at most NKC 64-deep chunks per
// pass, later passes accumulating into the output.
//
// Output slots: every slot of [start, roundup8(end)) is written (padding
// slots get 0), in the reference layout [(j/8)][v][j%8] (OUT_REF) or
// vector-major [j][v].
#pragma once

#include "mc_device.hpp"

namespace mc {

struct SddmmArgs {
    int m_vec;
    int k;
    int n;
    const int *row_indices;    // may be null (identity)
    const int *row_offsets;    // 2*m_vec (start, end)
    const int *col_indices;    // rhs row per slot, -1 in padding
    const unsigned char *lhs;
    const unsigned char *rhs;
    int *out;
    unsigned lhs_bytes;
    unsigned rhs_bytes;        // < 2^30
    // batched launch (blockIdx.y = matrix b of one shared mask): matrix b's
    // lhs / rhs start lhs_stride / rhs_stride bytes after matrix b-1's, its
    // output out_stride int32 after; 0 for a single matrix
    long long lhs_stride;
    long long rhs_stride;
    long long out_stride;
};

template <int BITS, bool SGN> struct SddmmDigits;
template <bool SGN> struct SddmmDigits<8, SGN> {
    static constexpr int P = SGN ? 1 : 2, OB = SGN ? 0 : 128;
    __device__ static constexpr int weight(int p) { return SGN ? 1 : (p == 0 ? 16 : 1); }
};
template <bool SGN> struct SddmmDigits<4, SGN> {
    static constexpr int P = 1, OB = SGN ? -8 : 0;
    __device__ static constexpr int weight(int) { return 1; }
};
template <bool SGN> struct SddmmDigits<16, SGN> {
    static constexpr int P = 2, OB = SGN ? 128 : 32896, OA = OB;
    __device__ static constexpr int weight(int p) { return p == 0 ? 256 : 1; }
};

// 16 k-values of 16-bit lanes at off_bytes as two byte-digit operands
// (k order kept): hi = high bytes, lo = low bytes, each XORed with its mask.
__device__ __forceinline__ void load_k16_w16(__amdgpu_buffer_rsrc_t r, u32 off_bytes, u32 xhi, u32 xlo, v4i &hi,
                                             v4i &lo) {
    const v4i w0 = buf_load_b128(r, off_bytes);
    const v4i w1 = buf_load_b128(r, off_bytes + 16u);
    const u32 w[8] = {(u32)w0[0], (u32)w0[1], (u32)w0[2], (u32)w0[3], (u32)w1[0], (u32)w1[1], (u32)w1[2], (u32)w1[3]};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        hi[q] = (int)(perm(w[2 * q + 1], w[2 * q], 0x07050301u) ^ xhi);
        lo[q] = (int)(perm(w[2 * q + 1], w[2 * q], 0x06040200u) ^ xlo);
    }
}

// 16 k-values of one packed row at k offset k0 (multiple of 16) as an int8
// operand.  8-bit: 16 bytes.  4-bit: 8 bytes expanded to [even nibbles(x0),
// odd(x0), even(x1), odd(x1)] — the same k order for both operands.
template <int BITS>
__device__ __forceinline__ v4i load_k16(__amdgpu_buffer_rsrc_t r, u32 off_bytes) {
    if constexpr (BITS == 8) {
        return buf_load_b128(r, off_bytes);
    } else {
        v2i x = buf_load_b64(r, off_bytes);
        return v4i{(int)x[0], (int)x[1], 0, 0};   // expanded by the caller
    }
}

__device__ __forceinline__ v4i expand_nibbles(v4i x, bool sext, u32 xor_mask) {
    u32 a = (u32)x[0] ^ xor_mask, b = (u32)x[1] ^ xor_mask;
    u32 e0 = nib_even(a), o0 = nib_odd(a), e1 = nib_even(b), o1 = nib_odd(b);
    if (sext) { e0 = sext4_bytes(e0); o0 = sext4_bytes(o0); e1 = sext4_bytes(e1); o1 = sext4_bytes(o1); }
    return v4i{(int)e0, (int)o0, (int)e1, (int)o1};
}

constexpr u32 kSddmmOob = 0x40000000u;

template <int V, int BITS, bool SGN, int NKC, int W, int TPW, bool OUT_REF>
__global__ __launch_bounds__(64 * W) void sddmm_kernel(SddmmArgs args) {
    using D = SddmmDigits<BITS, SGN>;
    constexpr int P = D::P;
    constexpr int ROWS = P * V;
    constexpr int OB = D::OB;
    constexpr bool W16 = BITS == 16;
    // 16-bit digit masks: unsigned bytes and signed low bytes become b ^ 0x80
    constexpr u32 XHI = SGN ? 0u : 0x80808080u, XLO = 0x80808080u;
    constexpr int HOME = V < 4 ? V : 4;
    constexpr int KB = 64 * BITS / 8;          // bytes of one 64-k chunk

    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int r = lane & 15, g = lane >> 4;
    const int m = uniform(args.row_indices ? args.row_indices[blockIdx.x] : (int)blockIdx.x);
    const int start = uniform(args.row_offsets[2 * m]);
    const int end = uniform(args.row_offsets[2 * m + 1]);
    const int wend = start + ((end - start + 7) & ~7);   // written slots

    const u32 row_bytes = (u32)args.k * BITS / 8;
    const long long bz = blockIdx.y;
    const __amdgpu_buffer_rsrc_t rl = make_rsrc(args.lhs + bz * args.lhs_stride, args.lhs_bytes);
    const __amdgpu_buffer_rsrc_t rr = make_rsrc(args.rhs + bz * args.rhs_stride, args.rhs_bytes);
    int *const out = args.out + bz * args.out_stride;

    // K runs in passes of up to NKC 64-deep chunks; every pass holds its A
    // operands in registers and adds its partial sums (exact mod 2^32, the
    // offset terms included) to the output slots the first pass wrote.
    const int chunks = args.k >> 6;
    const int passes = (chunks + NKC - 1) / NKC;
    const bool a_live = r < ROWS;
    const int a_p = r / V, a_v = r % V;
    for (int ps = 0; ps < passes; ++ps) {
        const int kc0 = ps * NKC;
        const int kn = min(NKC, chunks - kc0);   // chunks of this pass (uniform)

        // A operands (lhs rows of this vector row) for the pass's chunks, and
        // their digit row sums for the offset correction.
        v4i aop[NKC];
        int rsum = 0;
#pragma unroll
        for (int kc = 0; kc < NKC; ++kc) {
            const bool live = a_live && kc < kn;
            const u32 off =
                live ? (u32)(m * V + a_v) * row_bytes + (u32)((kc0 + kc) * KB + 16 * g * BITS / 8) : 0x80000000u;
            v4i x;
            if constexpr (W16) {
                v4i hi, lo;
                load_k16_w16(rl, off, XHI, XLO, hi, lo);
                // dead rows and chunks load nothing (offset out of range) but
                // keep the XOR masks: zero them so they sum to 0
                x = live ? (a_p == 0 ? hi : lo) : v4i{0, 0, 0, 0};
            } else {
                x = load_k16<BITS>(rl, off);
            }
            if constexpr (BITS == 8) {
                if constexpr (!SGN) {
                    const int sh = a_p == 0 ? 4 : 0;
#pragma unroll
                    for (int q = 0; q < 4; ++q) x[q] = (int)(((u32)x[q] >> sh) & 0x0F0F0F0Fu);
                }
            } else if constexpr (BITS == 4) {
                x = expand_nibbles(x, SGN, 0u);
            }
            aop[kc] = x;
            if constexpr (OB != 0) rsum += sum_bytes(x);
        }
        int rs_home[HOME];
#pragma unroll
        for (int i = 0; i < HOME; ++i) rs_home[i] = 0;
        if constexpr (OB != 0) {
            int s = rsum;
            s += lane_read(s, lane ^ 16);
            s += lane_read(s, lane ^ 32);   // full row sum of MFMA row r
#pragma unroll
            for (int i = 0; i < HOME; ++i) {
                const int v = 4 * g + i;
#pragma unroll
                for (int p = 0; p < P; ++p) rs_home[i] += D::weight(p) * lane_read(s, (p * V + v) & 15);
            }
        }

        {
            const int ctiles = (wend - start + 15) >> 4;
            const __amdgpu_buffer_rsrc_t rc = make_rsrc(args.col_indices + start, (u32)(end - start) * 4u);
            // wave w handles tile groups w, w + W, ... of TPW tiles each
            for (int tg = wave * TPW; tg < ctiles; tg += W * TPW) {
                constexpr int BD = W16 ? 2 : 1;   // rhs digits (16-bit: low, high)
                v4i bop[TPW][NKC][BD];
                int bsum[TPW];                    // sum of rhs' over this lane's k (16-bit)
                // each wave loads the column indices of its own tiles (lane r:
                // slot 16 tile + r) as rhs row byte offsets, out of range past
                // the row end and for tiles past the row's last
                u32 row_off[TPW];
#pragma unroll
                for (int t = 0; t < TPW; ++t) {
                    const int rel = (tg + t) * 16 + r;   // slot - start
                    const bool in = tg + t < ctiles && start + rel < end;
                    const int idx = (int)buf_load_b32(rc, in ? (u32)rel * 4u : 0x80000000u);
                    row_off[t] = in && idx >= 0 ? (u32)idx * row_bytes : kSddmmOob;
                }
#pragma unroll
                for (int t = 0; t < TPW; ++t) {
                    const u32 o = row_off[t];
                    bsum[t] = 0;
#pragma unroll
                    for (int kc = 0; kc < NKC; ++kc) {
                        // chunks past the pass read out of range (B < 1 GiB, so
                        // o + kSddmmOob stays below 2^32 and outside the buffer)
                        const u32 ko = o + (kc < kn ? (u32)((kc0 + kc) * KB) : kSddmmOob);
                        if constexpr (W16) {
                            v4i hi, lo;
                            load_k16_w16(rr, ko + (u32)(32 * g), XHI, XLO, hi, lo);
                            if (kc >= kn) { hi = v4i{0, 0, 0, 0}; lo = v4i{0, 0, 0, 0}; }
                            bop[t][kc][0] = lo;
                            bop[t][kc][BD - 1] = hi;
                            bsum[t] += 256 * sum_bytes(hi) + sum_bytes(lo);
                            continue;
                        }
                        v4i x = load_k16<BITS>(rr, ko + (u32)(16 * g * BITS / 8));
                        if constexpr (BITS == 8) {
                            if constexpr (!SGN) {
#pragma unroll
                                for (int q = 0; q < 4; ++q) x[q] = (int)((u32)x[q] ^ 0x80808080u);
                            }
                        } else {
                            x = expand_nibbles(x, false, SGN ? 0x88888888u : 0u);
                        }
                        bop[t][kc][0] = x;
                    }
                }
#pragma unroll
                for (int t = 0; t < TPW; ++t) {
                    v4i acc[BD];
#pragma unroll
                    for (int q = 0; q < BD; ++q) {
                        acc[q] = v4i{0, 0, 0, 0};
#pragma unroll
                        for (int kc = 0; kc < NKC; ++kc) acc[q] = mfma_i8(aop[kc], bop[t][kc][q], acc[q]);
                    }
                    // rhs digit sum of nonzero r over the pass's k (16-bit only)
                    u32 corr = 0;
                    if constexpr (W16) {
                        int sb = bsum[t];
                        sb += lane_read(sb, lane ^ 16);
                        sb += lane_read(sb, lane ^ 32);
                        corr = (u32)D::OA * (u32)sb + (u32)(64 * kn) * (u32)D::OA * (u32)OB;
                    }
                    // recombine digit planes into the home lanes (rows v = 4g + i),
                    // mod 2^32 in unsigned arithmetic
                    int res[HOME];
#pragma unroll
                    for (int i = 0; i < HOME; ++i) {
                        u32 vsum = (u32)OB * (u32)rs_home[i] + corr;
#pragma unroll
                        for (int p = 0; p < P; ++p) {
                            int dg, di;
                            if constexpr (V == 8) { dg = 2 * (p % 2); di = 0; }
                            else if constexpr (V == 4) { dg = p % 4; di = 0; }
                            else { dg = (p % 8) / 2; di = 2 * (p % 2); }
                            u32 x = (u32)acc[0][i + di];
                            if constexpr (W16) x += (u32)acc[1][i + di] << 8;
                            if (dg != 0) x = (u32)lane_down((int)x, 16 * dg);
                            vsum += (u32)D::weight(p) * x;
                        }
                        res[i] = (int)vsum;
                    }
                    const int j = start + (tg + t) * 16 + r;
                    // reference layout with four rows per lane: a 4 x 4 transpose
                    // within each quad of lanes gives lane r row 4g + (r & 3) of
                    // slots j4..j4+3 (j4 = j & ~3, start being a multiple of 8),
                    // one 16-byte run of [(j/8)][v][j%8].  A quad's four lanes are
                    // all in or all out of the store below (its slots share j4).
                    int tq[4] = {0, 0, 0, 0};
                    if constexpr (OUT_REF && HOME == 4) {
#pragma unroll
                        for (int i = 0; i < 4; ++i) tq[i] = j < end ? res[i] : 0;
                        quad_transpose4(tq, lane);
                    }
                    if (g < (V + 3) / 4 && tg + t < ctiles && j < wend) {
                        auto store_each = [&]() {
#pragma unroll
                            for (int i = 0; i < HOME; ++i) {
                                const int v = 4 * g + i;
                                const int val = j < end ? res[i] : 0;
                                const size_t o =
                                    OUT_REF ? (size_t)(j >> 3) * 8 * V + 8 * v + (j & 7) : (size_t)j * V + v;
                                // later passes add to what this lane stored before
                                out[o] = ps == 0 ? val : (int)((u32)out[o] + (u32)val);
                            }
                        };
                        if constexpr (OUT_REF && HOME == 4) {
                            if ((reinterpret_cast<uintptr_t>(out) & 15u) == 0) {
                                const int j4 = j & ~3;
                                const int v = 4 * g + (r & 3);
                                v4i val = v4i{tq[0], tq[1], tq[2], tq[3]};
                                v4i *o = reinterpret_cast<v4i *>(out + (size_t)(j4 >> 3) * 8 * V + 8 * v + (j4 & 7));
                                if (ps != 0) {
                                    const v4i prev = *o;
#pragma unroll
                                    for (int i = 0; i < 4; ++i) val[i] = (int)((u32)prev[i] + (u32)val[i]);
                                }
                                *o = val;
                            } else {
                                store_each();
                            }
                        } else if constexpr (!OUT_REF && HOME == 4) {
                            if ((reinterpret_cast<uintptr_t>(out) & 15u) == 0) {
                                // [slot][V]: the lane's rows 4g..4g+3 are one 16-byte run
                                v4i val = j < end ? v4i{res[0], res[1], res[2], res[3]} : v4i{0, 0, 0, 0};
                                v4i *o = reinterpret_cast<v4i *>(out + (size_t)j * V + 4 * g);
                                if (ps != 0) {
                                    const v4i prev = *o;
#pragma unroll
                                    for (int i = 0; i < 4; ++i) val[i] = (int)((u32)prev[i] + (u32)val[i]);
                                }
                                *o = val;
                            } else {
                                store_each();
                            }
                        } else {
                            store_each();
                        }
                    }
                }
            }
        }
    }
}

}  // namespace mc
