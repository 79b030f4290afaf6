// Quantized sparse attention on gfx950: quantize, softmax + requantize,
// dequantize, and the fused batched kernel (one wave per (matrix, vector row)
// work item, persistent over a longest-first work list), plus the C ABI.
#include <hip/hip_fp16.h>

#include "attention_common.hpp"
#include "mc_internal.hpp"
#include "sddmm_kernel.hpp"

namespace mc {

// Device view of mc_attn_matrix (same member order and types).
struct AttnMatrix {
    int L;
    int m_vec;
    const int *row_offsets;
    const int *col_indices;
    const unsigned char *q;
    const unsigned char *k;
    const unsigned char *v;
    float alpha;
    float o_scale;
    float *o;
    int *o_int;
    unsigned char *p_out;
    int *s_out;
    __half *o16;
    int o16_stride;
    int o16_pad;
    const __half *q16;
    int q16_stride;
    float q_scale;
    int q_rounding;
    int q_pad;
};
static_assert(sizeof(AttnMatrix) == sizeof(mc_attn_matrix), "descriptor layout");

// ---------------------------------------------------------------------------
// quantize: q = trunc(clamp(float(x) * scale, lo, hi)), or rint() for nearest
// (end2end_eval/sparse_transformer_magicube/src/cuda/quantization_kernel.cu:46-74
// for 8 bits, :13-43 for 4 bits; 4-bit lanes packed 8 per int32, lane 0 low)
// ---------------------------------------------------------------------------
struct QuantJob {
    const __half *x;
    long long count;
    float scale;
    void *out;
    long long row_len;      // 0: x is contiguous; else rows of row_len values,
    long long row_stride;   // row_stride values apart (row_len % 8 == 0)
};
static_assert(sizeof(QuantJob) == sizeof(mc_quant_job), "job layout");

// Groups of 8 consecutive values that one thread converts per pass: the
// block covers 256 * kQuantGroups groups, thread t takes groups t, t + 256,
// ... of them, and all of its loads are issued before any is converted
// (two 16-byte loads in flight per thread; one or four measured slower).
constexpr int kQuantGroups = 2;

template <int BITS>
__global__ __launch_bounds__(256) void quantize_kernel(const QuantJob *jobs, int njobs, int rounding) {
    const QuantJob jb = jobs[blockIdx.y];
    // MC_QUANT_RANGE4 (bit 1 of rounding): 4-bit values in int8 storage
    const bool range4 = BITS == 4 || (rounding & 2) != 0;
    const float lo = range4 ? -8.f : -128.f, hi = range4 ? 7.f : 127.f;
    constexpr int G = kQuantGroups;
    const long long groups = (jb.count + 7) / 8;
    for (long long blk = blockIdx.x; blk * 256 * G < groups; blk += gridDim.x) {
        const long long g0 = blk * 256 * G + threadIdx.x;
        v4i w[G];
        bool fast[G];
#pragma unroll
        for (int k = 0; k < G; ++k) {
            const long long gi = g0 + 256 * k;
            const long long base = gi * 8;
            // the 8 values never straddle a source row (row_len % 8 == 0)
            const __half *src = jb.row_len ? jb.x + (base / jb.row_len) * jb.row_stride + base % jb.row_len
                                           : jb.x + base;
            fast[k] = gi < groups && base + 8 <= jb.count && (reinterpret_cast<uintptr_t>(src) & 15u) == 0;
            // one 16-byte load for the 8 values, non-temporal: the fp16 source is
            // read once, and keeping it out of the caches leaves them to the
            // quantized operands the fused kernel gathers next
            if (fast[k]) w[k] = __builtin_nontemporal_load(reinterpret_cast<const v4i *>(src));
        }
#pragma unroll
        for (int k = 0; k < G; ++k) {
            const long long gi = g0 + 256 * k;
            if (gi >= groups) continue;
            const long long base = gi * 8;
            float x[8];
            if (fast[k]) {
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int we = w[k][e];
                    const __half2 h2 = *reinterpret_cast<const __half2 *>(&we);
                    x[2 * e] = __low2float(h2);
                    x[2 * e + 1] = __high2float(h2);
                }
            } else {
                const __half *src = jb.row_len ? jb.x + (base / jb.row_len) * jb.row_stride + base % jb.row_len
                                               : jb.x + base;
#pragma unroll
                for (int e = 0; e < 8; ++e) x[e] = base + e < jb.count ? __half2float(src[e]) : 0.f;
            }
            int q[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                float f = x[e] * jb.scale;
                f = fminf(fmaxf(f, lo), hi);
                q[e] = (rounding & 1) == 0 ? (int)f : (int)rintf(f);
            }
            if constexpr (BITS == 8) {
                unsigned char *o = static_cast<unsigned char *>(jb.out) + base;
                if (base + 8 <= jb.count) {
                    u32 w0 = (u32)(q[0] & 255) | (u32)(q[1] & 255) << 8 | (u32)(q[2] & 255) << 16 | (u32)(q[3] & 255) << 24;
                    u32 w1 = (u32)(q[4] & 255) | (u32)(q[5] & 255) << 8 | (u32)(q[6] & 255) << 16 | (u32)(q[7] & 255) << 24;
                    *reinterpret_cast<v2i *>(o) = v2i{(int)w0, (int)w1};
                } else {
                    for (int e = 0; base + e < jb.count; ++e) o[e] = (unsigned char)q[e];
                }
            } else {
                u32 wq = 0;
#pragma unroll
                for (int e = 0; e < 8; ++e) wq |= (u32)(q[e] & 15) << (4 * e);
                static_cast<u32 *>(jb.out)[gi] = wq;
            }
        }
    }
    (void)njobs;
}

// max |x| over a tensor, combined with atomicMax on the float bit pattern
// (non-negative floats order like their bits); *out must start at 0.
__global__ __launch_bounds__(256) void absmax_kernel(const __half *x, long long count, unsigned *out) {
    float m = 0.f;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < count; i += (long long)gridDim.x * 256)
        m = fmaxf(m, fabsf(__half2float(x[i])));
    for (int s = 32; s > 0; s >>= 1) m = fmaxf(m, __shfl_xor(m, s, 64));
    __shared__ float part[4];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        const float b = fmaxf(fmaxf(part[0], part[1]), fmaxf(part[2], part[3]));
        atomicMax(out, __float_as_uint(b));
    }
}

// dequantize: out[r][c] = (float(in[r][c]) * scale) * row_scale[r]  (row_scale optional)
__global__ __launch_bounds__(256) void dequant_kernel(const int *in, long long count, int cols, float scale,
                                                      const float *row_scale, float *out) {
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < count; i += (long long)gridDim.x * 256) {
        float f = (float)in[i] * scale;
        if (row_scale) f = f * row_scale[i / cols];
        out[i] = f;
    }
}

// ---------------------------------------------------------------------------
// softmax + requantize (q_softmax_kernel.cu:180-271): one wave per vector
// row.  S is int32 vector-major [slot][V] (the SDDMM output); P is written
// in the SpMM A layout: groups of 16 slots, [V][16] bytes per group.
// ---------------------------------------------------------------------------
template <int V, int WPB>
__global__ __launch_bounds__(64 * WPB) void softmax_requant_kernel(int m_vec, const int *row_indices,
                                                                   const int *row_offsets, const int *s_vals,
                                                                   float alpha, float p_scale, int p_mode,
                                                                   unsigned char *p_out, float *row_inv_l) {
    constexpr int HOME = V < 4 ? V : 4;
    __shared__ __attribute__((aligned(16))) unsigned char tiles[WPB][16 * 64];
    // the wave id in an SGPR keeps work-item indices, descriptors and buffer
    // resources wave-uniform (no per-lane waterfall around each buffer load)
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int row_slot = blockIdx.x * WPB + wave;
    if (row_slot >= m_vec) return;
    const int c = lane & 15, g = lane >> 4;
    const bool home = g < (V + 3) / 4;
    const int i_row = uniform(row_indices ? row_indices[row_slot] : row_slot);
    const int start = uniform(row_offsets[2 * i_row]), end = uniform(row_offsets[2 * i_row + 1]);
    const int nch = (end - start + 63) >> 6;

    auto load_chunk = [&](int ch, float (&s)[4][4]) {
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int slot = start + 64 * ch + 16 * t + c;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int v = 4 * g + i;
                s[t][i] = (home && i < HOME && slot < end) ? (float)s_vals[(size_t)slot * V + v] * alpha : -INFINITY;
            }
        }
    };

    float m[HOME], l[HOME];
#pragma unroll
    for (int i = 0; i < HOME; ++i) { m[i] = -INFINITY; l[i] = 0.f; }
    for (int ch = 0; ch < nch; ++ch) {
        float s[4][4];
        load_chunk(ch, s);
        softmax_chunk_update<HOME>(s, m, l);
    }
    float inv_l[HOME];
#pragma unroll
    for (int i = 0; i < HOME; ++i) inv_l[i] = l[i] > 0.f ? 1.0f / l[i] : 0.f;
    float mq[HOME];
#pragma unroll
    for (int i = 0; i < HOME; ++i) mq[i] = finite_row_max(m[i]);
    if (row_inv_l && home && c == 0) {
#pragma unroll
        for (int i = 0; i < HOME; ++i) row_inv_l[(size_t)i_row * V + 4 * g + i] = inv_l[i];
    }
    // padded slots of the row (up to the next multiple of 16) get P = 0
    const int aend = start + ((end - start + 15) & ~15);
    const int nch_w = (aend - start + 63) >> 6;
    for (int ch = 0; ch < nch_w; ++ch) {
        float s[4][4];
        load_chunk(ch, s);
        int pq[4][4];
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int i = 0; i < 4; ++i)
                pq[t][i] = i < HOME ? quantize_prob(s[t][i], mq[i < HOME ? i : 0], inv_l[i < HOME ? i : 0], p_scale, p_mode) : 0;
        const v4i a = p_to_a_layout<HOME>(tiles[wave], pq, lane, home, 4 * g);
        const int slot0 = start + 64 * ch + 16 * g;        // first slot of this lane's 16-slot group
        if (c < V && slot0 < aend)
            *reinterpret_cast<v4i *>(p_out + (size_t)(slot0 >> 4) * 16 * V + 16 * c) = a;
    }
}

// dequantize to fp16: out[i] = half(float(in[i]) / divisor), round to nearest
// even (the reference's epilogues, deq_spmm_kernel.cu:645-670 and
// sddmm_utils/output_tile.h:37-43: __float2half((float)acc / scale))
__global__ __launch_bounds__(256) void dequant_f16_kernel(const int *in, long long count, float divisor,
                                                          __half *out) {
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < count; i += (long long)gridDim.x * 256)
        out[i] = __float2half_rn((float)in[i] / divisor);
}

// ---------------------------------------------------------------------------
// The reference's batched quantized softmax on fp16 logits
// (q_softmax_kernel.cu:180-271, host :394-479): for matrix b of a batch
// sharing one mask and each vector row m = row_indices[x],
//   s = float(values[slot][v]) * sqrt_dk
//   max_v = max over the row's nonzeros of s, starting from -1e10 for v = 0
//           and from 0 for v >= 1 (the reference's register init sets only
//           element 0, :209)
//   sum_v = sum of exp(s - max_v)
//   P = (char / short) clamp(exp(s - max_v) / sum_v * scale, lo, hi)
// written at the reference's layouts (attn + start * V * bits / 8 bytes):
//   8-bit   [(i/16)][v][16] bytes
//   16-bit  [(i/16)][v][low 16 bytes][high 16 bytes]
//   4-bit   [(i/32)][v][16 bytes], byte j = slot 2j (low nibble) | 2j+1
// Every slot of [start, start + roundup(nonzeros, A)) is written (A = 32 for
// 4-bit P, else 16), padding slots with 0.  The sum runs in a fixed order:
// lane l adds slots l, l + 64, ..., then a butterfly over the 64 lanes.
// ---------------------------------------------------------------------------
template <int V>
__device__ __forceinline__ void load_half_vec(const __half *p, float (&x)[V]) {
    if constexpr (V == 8) {
        const v4i w = *reinterpret_cast<const v4i *>(p);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int we = w[e];
            const __half2 h2 = *reinterpret_cast<const __half2 *>(&we);
            x[2 * e] = __low2float(h2);
            x[2 * e + 1] = __high2float(h2);
        }
    } else {
#pragma unroll
        for (int v = 0; v < V; ++v) x[v] = __half2float(p[v]);
    }
}

__device__ __forceinline__ float wave_max(float x) {
    for (int s = 32; s > 0; s >>= 1) x = fmaxf(x, __shfl_xor(x, s, 64));
    return x;
}
__device__ __forceinline__ float wave_sum(float x) {
    for (int s = 1; s < 64; s <<= 1) x += __shfl_xor(x, s, 64);
    return x;
}

template <int V, int BITS, int WPB>
__global__ __launch_bounds__(64 * WPB) void ref_softmax_kernel(int rows, const int *row_indices, const int *row_offsets,
                                                               const __half *values, long long values_stride,
                                                               float sqrt_dk, float scale, unsigned char *attn,
                                                               long long attn_stride) {
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int x = blockIdx.x * WPB + wave;
    if (x >= rows) return;
    const long long b = blockIdx.y;
    const int m = uniform(row_indices ? row_indices[x] : x);
    const int start = uniform(row_offsets[2 * m]), nz = uniform(row_offsets[2 * m + 1]) - start;
    const __half *val = values + b * values_stride + (size_t)start * V;
    unsigned char *out = attn + b * attn_stride + (size_t)start * V * BITS / 8;

    float mx[V], sm[V];
#pragma unroll
    for (int v = 0; v < V; ++v) { mx[v] = v == 0 ? -1e10f : 0.f; sm[v] = 0.f; }
    for (int i = lane; i < nz; i += 64) {
        float s[V];
        load_half_vec<V>(val + (size_t)i * V, s);
#pragma unroll
        for (int v = 0; v < V; ++v) mx[v] = fmaxf(mx[v], s[v] * sqrt_dk);
    }
#pragma unroll
    for (int v = 0; v < V; ++v) mx[v] = wave_max(mx[v]);
    for (int i = lane; i < nz; i += 64) {
        float s[V];
        load_half_vec<V>(val + (size_t)i * V, s);
#pragma unroll
        for (int v = 0; v < V; ++v) sm[v] += __expf(s[v] * sqrt_dk - mx[v]);
    }
#pragma unroll
    for (int v = 0; v < V; ++v) sm[v] = wave_sum(sm[v]);

    auto prob = [&](int i, int v, const float (&s)[V], float lo, float hi) {
        if (i >= nz) return 0.f;
        const float p = __expf(s[v] * sqrt_dk - mx[v]) / sm[v] * scale;
        return fminf(fmaxf(p, lo), hi);
    };
    if constexpr (BITS == 4) {
        // slot pairs (2j, 2j + 1) of 32-slot groups
        const int npair = ((nz + 31) & ~31) / 2;
        for (int j = lane; j < npair; j += 64) {
            const int i = 2 * j;
            float s0[V], s1[V];
            if (i < nz) load_half_vec<V>(val + (size_t)i * V, s0);
            if (i + 1 < nz) load_half_vec<V>(val + (size_t)(i + 1) * V, s1);
#pragma unroll
            for (int v = 0; v < V; ++v) {
                const int p0 = (int)(signed char)prob(i, v, s0, -8.f, 7.f);
                const int p1 = (int)(signed char)prob(i + 1, v, s1, -8.f, 7.f);
                out[(size_t)(i >> 5) * V * 16 + 16 * v + (j & 15)] = (unsigned char)((p0 & 15) | (p1 << 4));
            }
        }
    } else {
        const int nw = (nz + 15) & ~15;
        for (int i = lane; i < nw; i += 64) {
            float s[V];
            if (i < nz) load_half_vec<V>(val + (size_t)i * V, s);
#pragma unroll
            for (int v = 0; v < V; ++v) {
                if constexpr (BITS == 8) {
                    const int p = (int)(signed char)prob(i, v, s, -128.f, 127.f);
                    out[(size_t)(i >> 4) * V * 16 + 16 * v + (i & 15)] = (unsigned char)p;
                } else {
                    const int p = (int)(short)prob(i, v, s, -32768.f, 32767.f);
                    unsigned char *o = out + (size_t)(i >> 4) * V * 32 + 32 * v + (i & 15);
                    o[0] = (unsigned char)(p & 255);
                    o[16] = (unsigned char)((p >> 8) & 255);
                }
            }
        }
    }
}

// Row bounds are read-only for the whole launch; read through the constant
// address space at a wave-uniform address they become scalar loads (scalar
// cache, lgkmcnt), which neither wait behind nor hold back vector gathers.
__device__ __forceinline__ int const_load(const int *p) {
    typedef const __attribute__((address_space(4))) int const_int;
    return *(const const_int *)p;
}

// ---------------------------------------------------------------------------
// Fused batched attention: persistent waves walk a work list of (matrix,
// vector row) items.  Pass 1 gathers K rows, computes the int32 logits with
// MFMA and keeps the online softmax statistics; pass 2 recomputes the exact
// logits, quantizes P, moves it to the A layout through LDS, gathers V rows
// (4x4 byte transposes) and accumulates O = P V with MFMA.
// VB = 4: V rows are packed 4-bit ([L][32] bytes, lane 0 low); each lane
// gathers the 16 bits of its four columns and widens them to the bytes
// V + 8 (0..15), so the MFMA forms P (V + 8) and 8 * (row sum of P) comes off
// the accumulators at the end (exact mod 2^32).
// ---------------------------------------------------------------------------
// DBG instantiations also write the optional int32 S, int8 P and int32 P V
// outputs; the production ones carry no stores inside the chunk loops (a
// conditional store there would make every load wait drain the queue).
// Pass 2 takes the scaled logits of a row's first chunks from an LDS copy
// made in pass 1 instead of gathering K and recomputing them: 12 KB per wave
// (8 floats per lane and chunk at V = 8, so six chunks; 4 and twelve at
// V = 2 and 4), 52 KB per workgroup with the P tiles, so three workgroups
// (three waves per SIMD, at most 168 registers) fit a CU's 160 KB.  Measured
// on 128 matrices of L ~ U[512, 4096]: 0.199 ms per step, against 0.233 ms
// with no copy at four waves per SIMD, 0.209 ms with a copy of three chunks
// of K instead, 0.210 ms with 8 KB of logits at four waves.
constexpr int kLogitStashBytes = 12288;   // per wave

template <int V, int QKB, int VB, int PM, int WPB, bool DBG>
__global__ __launch_bounds__(64 * WPB) __attribute__((amdgpu_waves_per_eu(3))) void attention_kernel(const AttnMatrix *mats, const int2 *items, int n_items,
                                                             float p_scale, int p_mode) {
    constexpr int HOME = V < 4 ? V : 4;
    constexpr int QROW = kHeadDim * QKB / 8;          // bytes per Q/K row
    constexpr int KCB = QROW / 4;                     // bytes of a K row per lane
    __shared__ __attribute__((aligned(16))) unsigned char tiles[WPB][16 * 64];
    // logits s[t][j] of a row's first SST chunks, kept from pass 1 for pass 2
    // (per wave; SQ 16-byte slots per lane, lane-major, so the accesses are
    // conflict-free)
    constexpr int HC_ = V == 8 ? 2 : 1;
    constexpr int SQ = HC_;                       // 4 * HC floats per lane
    constexpr int SST = kLogitStashBytes / (64 * 16 * SQ);
    __shared__ v4f sstash[WPB][SST][SQ][64];
    // the wave id in an SGPR keeps work-item indices, descriptors and buffer
    // resources wave-uniform (no per-lane waterfall around each buffer load)
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int c = lane & 15, g = lane >> 4;
    const bool home = g < (V + 3) / 4;
    // Softmax rows per lane: the MFMA C tile holds the V real rows in lane
    // groups g < (V + 3) / 4 (rows 4g + i); permlane swaps deal them to all
    // four groups first, so that the per-logit work runs on real rows only.
    // V = 8: group g holds rows row0 + {0, 1}, row0 = 4 (g & 1) + 2 (g >> 1);
    // V = 4: group g holds row g; V = 2: groups 0 and 1 hold rows 0 and 1.
    constexpr int HC = V == 8 ? 2 : 1;
    const bool chome = V == 2 ? g < 2 : true;
    const int row0 = V == 8 ? 4 * (g & 1) + 2 * (g >> 1) : g;
    const int nwaves = gridDim.x * WPB;

    for (int it = blockIdx.x * WPB + wave; it < n_items; it += nwaves) {
        const int2 item = items[it];
        if (uniform(item.x) < 0) continue;           // padding slot of the work list
        const AttnMatrix Mv = mats[uniform(item.x)];
        AttnMatrix M;
        M.L = uniform(Mv.L);
        M.m_vec = uniform(Mv.m_vec);
        M.row_offsets = uniform(Mv.row_offsets);
        M.col_indices = uniform(Mv.col_indices);
        M.q = uniform(Mv.q);
        M.k = uniform(Mv.k);
        M.v = uniform(Mv.v);
        M.alpha = uniform(Mv.alpha);
        M.o_scale = uniform(Mv.o_scale);
        M.o = uniform(Mv.o);
        M.o_int = uniform(Mv.o_int);
        M.p_out = uniform(Mv.p_out);
        M.s_out = uniform(Mv.s_out);
        M.o16 = uniform(Mv.o16);
        M.o16_stride = uniform(Mv.o16_stride);
        M.q16 = uniform(Mv.q16);
        M.q16_stride = uniform(Mv.q16_stride);
        M.q_scale = uniform(Mv.q_scale);
        M.q_rounding = uniform(Mv.q_rounding);
        const int i_row = uniform(item.y);
        const int start = uniform(const_load(M.row_offsets + 2 * i_row)),
                  end = uniform(const_load(M.row_offsets + 2 * i_row + 1));
        const int aend = start + ((end - start + 15) & ~15);
        const int nch = (end - start + 63) >> 6;
        const int nch_w = (aend - start + 63) >> 6;
        const u32 kv_bytes = (u32)M.L * QROW;
        const __amdgpu_buffer_rsrc_t rk = make_rsrc(M.k, kv_bytes);
        const __amdgpu_buffer_rsrc_t rv = make_rsrc(M.v, (u32)M.L * (kHeadDim * VB / 8));
        const __amdgpu_buffer_rsrc_t rq = make_rsrc(M.q, kv_bytes);

        // SDDMM A operand: Q rows of this vector row (rows >= V read out of range -> 0)
        v4i qop = v4i{0, 0, 0, 0};
        // fp16 Q (q16 non-null): lane (c, g) loads columns 16g .. 16g + 15
        // of row i_row * V + c here and quantizes them once the pipeline's
        // first index and K loads are out (quantize_q16)
        v4i qh0, qh1;
        if (M.q16) {
            const __amdgpu_buffer_rsrc_t rq16 =
                make_rsrc(M.q16, (u32)(((u32)(M.L - 1) * (u32)M.q16_stride + kHeadDim) * 2u));
            const u32 off = c < V ? ((u32)(i_row * V + c) * (u32)M.q16_stride + (u32)(16 * g)) * 2u : 0x80000000u;
            qh0 = buf_load_b128(rq16, off);
            qh1 = buf_load_b128(rq16, off + 16u);
        } else {
            const u32 off = c < V ? (u32)(i_row * V + c) * QROW + (u32)(16 * g * QKB / 8) : 0x80000000u;
            qop = load_k16<QKB>(rq, off);
            if constexpr (QKB == 4) qop = expand_nibbles(qop, true, 0u);
        }
        // the quantize launch's expression (quantize_kernel), laid out as the
        // quantized operand would arrive: in order for int8; for int4 in
        // expand_nibbles' order (even, then odd columns of each 8)
        auto quantize_q16 = [&]() {
            const bool range4 = QKB == 4 || (M.q_rounding & 2) != 0;
            const float lo = range4 ? -8.f : -128.f, hi = range4 ? 7.f : 127.f;
            u32 qb[16];
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                const int w = e < 8 ? qh0[e >> 1] : qh1[(e - 8) >> 1];
                const __half2 h2 = *reinterpret_cast<const __half2 *>(&w);
                float f = ((e & 1) ? __high2float(h2) : __low2float(h2)) * M.q_scale;
                f = fminf(fmaxf(f, lo), hi);
                qb[e] = (u32)((M.q_rounding & 1) == 0 ? (int)f : (int)rintf(f)) & 255u;
            }
            const auto pack = [&](int a, int b, int c2, int d) {
                return (int)(qb[a] | qb[b] << 8 | qb[c2] << 16 | qb[d] << 24);
            };
            if constexpr (QKB == 8) qop = v4i{pack(0, 1, 2, 3), pack(4, 5, 6, 7), pack(8, 9, 10, 11), pack(12, 13, 14, 15)};
            else qop = v4i{pack(0, 2, 4, 6), pack(1, 3, 5, 7), pack(8, 10, 12, 14), pack(9, 11, 13, 15)};
        };

        // ---- software pipeline over 64-slot chunks.  Per half step for chunk
        // ch: issue the index load of chunk ch+2, the row gathers of chunk ch+1
        // (its index arrived one half step ago), then consume chunk ch.  Loads
        // retire in order, so each consumer waits only for older loads and the
        // gathers of the next chunk stay in flight.  Chunks past the row end
        // are fully masked: they leave (m, l) bitwise unchanged and add P = 0,
        // so the loops run in pairs without branches.
        const __amdgpu_buffer_rsrc_t rc = make_rsrc(M.col_indices, 0x7FFFFFFFu);
        // per lane: byte offset of its slot's V row (index * 64), or the
        // out-of-range marker for slots past the row end (L * 64 < 2^31)
        constexpr u32 kRowOob = 0x80000000u;
        auto load_idx = [&](int ch) {
            const int slot = start + 64 * ch + lane;
            const u32 v = buf_load_b32(rc, slot < end ? (u32)slot * 4u : kRowOob);
            return slot < end ? v * (u32)kHeadDim : kRowOob;
        };
        // K rows of slots 16t + c, bytes [KCB * g, +KCB): the MFMA B operand
        // directly (gathering four lanes per row and moving the rows to this
        // layout through LDS makes fewer L1 accesses but measured slower)
        auto load_krows = [&](u32 my_off, v4i (&kop)[4]) {
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                u32 row = (u32)__shfl((int)my_off, 16 * t + c, 64);
                if constexpr (QKB == 4) row = (row >> 1) | (row & kRowOob);   // 32-byte rows, marker kept
                kop[t] = load_k16<QKB>(rk, row + (u32)(KCB * g));
            }
        };
        // V rows of slots 16g + p (p = 0..15), columns 4c..4c+3: dword c of an
        // int8 row, or the 16 bits at byte 2c of a packed 4-bit row
        auto load_vrows = [&](u32 my_off, u32 (&raw)[16]) {
#pragma unroll
            for (int p = 0; p < 16; ++p) {
                u32 row = (u32)row_bcast16((int)my_off, p);   // lane 16g + p
                if constexpr (VB == 4) {
                    row = (row >> 1) | (row & kRowOob);   // 32-byte rows, marker kept
                    raw[p] = buf_load_u16(rv, row + 2u * c);
                } else {
                    raw[p] = buf_load_b32(rv, row + 4u * c);
                }
            }
        };
        // C-tile rows 4g + i of one 16-slot tile -> this lane's rows row0 + j
        auto compact_rows = [&](const v4i &acc, int (&out)[HC]) {
            if constexpr (V == 8) {
                // lanes 32..63 take rows 4g' + 2, 4g' + 3 of lanes 0..31 (g' = g - 2)
                out[0] = (int)__builtin_amdgcn_permlane32_swap((u32)acc[0], (u32)acc[2], false, false)[0];
                out[1] = (int)__builtin_amdgcn_permlane32_swap((u32)acc[1], (u32)acc[3], false, false)[0];
            } else if constexpr (V == 4) {
                // rows 0, 1 to groups 0, 1 and rows 2, 3 to groups 2, 3
                const u32 a = __builtin_amdgcn_permlane16_swap((u32)acc[0], (u32)acc[1], false, false)[0];
                const u32 b = __builtin_amdgcn_permlane16_swap((u32)acc[2], (u32)acc[3], false, false)[0];
                out[0] = (int)__builtin_amdgcn_permlane32_swap(a, b, false, false)[0];
            } else {
                out[0] = (int)__builtin_amdgcn_permlane16_swap((u32)acc[0], (u32)acc[1], false, false)[0];
            }
        };
        // logits of one chunk, scaled, in this lane's rows; masked -> -inf.
        // si: the raw C tile (rows 4g + i), kept for the debug outputs only
        auto logits = [&](int ch, const v4i (&kop)[4], float (&s)[4][HC], int (&si)[4][4]) {
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                v4i kk = kop[t];
                if constexpr (QKB == 4) kk = expand_nibbles(kk, true, 0u);
                const v4i acc = mfma_i8(qop, kk, v4i{0, 0, 0, 0});
                const int slot = start + 64 * ch + 16 * t + c;
                if constexpr (DBG) {
#pragma unroll
                    for (int i = 0; i < 4; ++i) si[t][i] = acc[i];
                }
                int v[HC];
                compact_rows(acc, v);
#pragma unroll
                for (int j = 0; j < HC; ++j) s[t][j] = (chome && slot < end) ? (float)v[j] * M.alpha : -INFINITY;
            }
        };

        float m[HC], l[HC];
#pragma unroll
        for (int j = 0; j < HC; ++j) { m[j] = -INFINITY; l[j] = 0.f; }
        auto pass1_chunk = [&](int ch, const v4i (&kop)[4]) {
            float s[4][HC];
            int si[4][4];
            logits(ch, kop, s, si);
            if (ch < SST) {
#pragma unroll
                for (int q = 0; q < SQ; ++q) {
                    v4f x;
#pragma unroll
                    for (int e = 0; e < 4; ++e) x[e] = s[(4 * q + e) / HC][(4 * q + e) % HC];
                    sstash[wave][ch][q][lane] = x;
                }
            }
            softmax_chunk_update<HC>(s, m, l);
            if (DBG && M.s_out && home) {
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const int slot = start + 64 * ch + 16 * t + c;
#pragma unroll
                    for (int i = 0; i < HOME; ++i)
                        if (slot < end) M.s_out[(size_t)slot * V + 4 * g + i] = si[t][i];
                }
            }
        };
        {
            v4i ka[4], kb[4];
            u32 ia = load_idx(0);
            u32 ib = load_idx(1);
            load_krows(ia, ka);
            if (M.q16) quantize_q16();
            for (int ch = 0; ch < nch; ch += 2) {
                ia = load_idx(ch + 2);
                load_krows(ib, kb);
                pass1_chunk(ch, ka);
                __builtin_amdgcn_sched_barrier(0);
                ib = load_idx(ch + 3);
                load_krows(ia, ka);
                pass1_chunk(ch + 1, kb);
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        float inv_l[HC];
#pragma unroll
        for (int j = 0; j < HC; ++j) inv_l[j] = l[j] > 0.f ? 1.0f / l[j] : 0.f;
        float mq[HC];
#pragma unroll
        for (int j = 0; j < HC; ++j) mq[j] = finite_row_max(m[j]);

        v4i oacc[4];
#pragma unroll
        for (int b = 0; b < 4; ++b) oacc[b] = v4i{0, 0, 0, 0};
        int psum = 0;   // VB = 4: sum of P row c over this lane's slots 16g..16g+15 of every chunk
        auto pass2_chunk = [&](int ch, const v4i (&kop)[4], const u32 (&raw_in)[16]) {
            float s[4][HC];
            int si[4][4];
            if (ch < SST) {
#pragma unroll
                for (int q = 0; q < SQ; ++q) {
                    const v4f x = sstash[wave][ch][q][lane];
#pragma unroll
                    for (int e = 0; e < 4; ++e) s[(4 * q + e) / HC][(4 * q + e) % HC] = x[e];
                }
            } else {
                logits(ch, kop, s, si);
            }
            int pq[4][HC];
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int j = 0; j < HC; ++j) pq[t][j] = quantize_prob<PM>(s[t][j], mq[j], inv_l[j], p_scale, p_mode);
            const v4i pop = p_to_a_layout<HC>(tiles[wave], pq, lane, chome, row0);
            u32 raw[16];
#pragma unroll
            for (int p = 0; p < 16; ++p) {
                if constexpr (VB == 4) {
                    // nibbles n0..n3 of V -> bytes n0 + 8 .. n3 + 8
                    const u32 x = raw_in[p] ^ 0x8888u;
                    raw[p] = perm((x >> 4) & 0x0F0Fu, x & 0x0F0Fu, 0x05010400u);
                } else {
                    raw[p] = raw_in[p];
                }
            }
            if constexpr (VB == 4) psum += sum_bytes(pop);
            if (DBG && M.p_out) {
                const int slot0 = start + 64 * ch + 16 * g;
                if (c < V && slot0 < aend)
                    *reinterpret_cast<v4i *>(M.p_out + (size_t)(slot0 >> 4) * 16 * V + 16 * c) = pop;
            }
            // after the 4x4 byte transposes, tr[q][b] holds column 4c + b of rows 4q..4q+3
            u32 tr[4][4];
#pragma unroll
            for (int q = 0; q < 4; ++q) transpose4x4(raw[4 * q], raw[4 * q + 1], raw[4 * q + 2], raw[4 * q + 3], tr[q]);
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const v4i bop = v4i{(int)tr[0][b], (int)tr[1][b], (int)tr[2][b], (int)tr[3][b]};
                oacc[b] = mfma_i8(pop, bop, oacc[b]);
            }
        };
        {
            // the first SST chunks' logits come from pass 1's LDS copy (the
            // same values pass 2 would recompute), the rest from K gathered
            // again
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            auto get_krows = [&](int ch, u32 my_off, v4i (&kop)[4]) {
                if (ch >= SST) load_krows(my_off, kop);
            };
            v4i ka[4], kb[4];
            u32 va[16], vb[16];
            u32 ia = load_idx(0);
            u32 ib = load_idx(1);
            get_krows(0, ia, ka);
            load_vrows(ia, va);
            for (int ch = 0; ch < nch_w; ch += 2) {
                ia = load_idx(ch + 2);
                get_krows(ch + 1, ib, kb);
                load_vrows(ib, vb);
                pass2_chunk(ch, ka, va);
                __builtin_amdgcn_sched_barrier(0);
                ib = load_idx(ch + 3);
                get_krows(ch + 2, ia, ka);
                load_vrows(ia, va);
                pass2_chunk(ch + 1, kb, vb);
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        // 1 / row sum of output rows 4g + i, from the lanes that hold those rows
        float row_inv_l[HOME];
#pragma unroll
        for (int i = 0; i < HOME; ++i) {
            const int src_group = V == 8 ? (g | (i & 2)) : ((4 * g + i) & 3);
            row_inv_l[i] = __shfl(inv_l[V == 8 ? (i & 1) : 0], 16 * src_group + c, 64);
        }
        if constexpr (VB == 4) {
            // row c's sum of P over all slots (the four lane groups hold
            // slots 16g..16g+15 of every chunk), then 8 * sum(P) of output
            // row 4g + i off its accumulators
            int s = psum;
            s += lane_read(s, lane ^ 16);
            s += lane_read(s, lane ^ 32);
#pragma unroll
            for (int i = 0; i < HOME; ++i) {
                const u32 corr = 8u * (u32)lane_read(s, 4 * g + i);
#pragma unroll
                for (int b = 0; b < 4; ++b) oacc[b][i] = (int)((u32)oacc[b][i] - corr);
            }
        }
        // rows 4g+i, columns 4c + b
        if (home) {
#pragma unroll
            for (int i = 0; i < HOME; ++i) {
                const size_t orow = (size_t)(i_row * V + 4 * g + i) * kHeadDim + 4 * c;
                if (DBG && M.o_int)
                    *reinterpret_cast<v4i *>(M.o_int + orow) = v4i{oacc[0][i], oacc[1][i], oacc[2][i], oacc[3][i]};
                // same float operations as dequant_kernel with/without a row scale
                const float rs = PM == kPModeRowScaled ? row_inv_l[i] : 1.f;
                float4 of;
                of.x = ((float)oacc[0][i] * M.o_scale) * rs;
                of.y = ((float)oacc[1][i] * M.o_scale) * rs;
                of.z = ((float)oacc[2][i] * M.o_scale) * rs;
                of.w = ((float)oacc[3][i] * M.o_scale) * rs;
                if (M.o16) {
                    // fp16 (round to nearest even) straight into a strided
                    // [L][row stride] layout, e.g. the [L][batch][heads * 64]
                    // input of the output projection
                    const __half2 h01 = __floats2half2_rn(of.x, of.y), h23 = __floats2half2_rn(of.z, of.w);
                    const size_t hrow = (size_t)(i_row * V + 4 * g + i) * (size_t)M.o16_stride + 4 * c;
                    *reinterpret_cast<v2i *>(M.o16 + hrow) =
                        v2i{*reinterpret_cast<const int *>(&h01), *reinterpret_cast<const int *>(&h23)};
                } else {
                    *reinterpret_cast<float4 *>(M.o + orow) = of;
                }
            }
        }
    }
}

}  // namespace mc

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
namespace {

inline unsigned grid_for(long long work, long long per_block) {
    long long b = (work + per_block - 1) / per_block;
    if (b < 1) b = 1;
    if (b > 65535) b = 65535;
    return (unsigned)b;
}

}  // namespace

extern "C" {

mc_status mc_attn_quantize_batched(const mc_quant_job *jobs, int njobs, int bits, int rounding,
                                   long long max_count, hipStream_t stream) {
    if (bits != 8 && bits != 4) return MC_EUNSUPPORTED;
    if ((rounding & ~MC_QUANT_RANGE4) != MC_ROUND_TRUNC && (rounding & ~MC_QUANT_RANGE4) != MC_ROUND_NEAREST)
        return MC_EINVAL_ARG;
    if (njobs < 0 || max_count < 0) return MC_EINVAL_ARG;
    if (njobs == 0 || max_count == 0) return MC_OK;
    if (!jobs || njobs > 65535) return MC_EINVAL_ARG;
    const unsigned gx = grid_for((max_count + 7) / 8, 256 * mc::kQuantGroups);
    const auto *dj = reinterpret_cast<const mc::QuantJob *>(jobs);
    if (bits == 8) hipLaunchKernelGGL(mc::quantize_kernel<8>, dim3(gx, njobs), dim3(256), 0, stream, dj, njobs, rounding);
    else hipLaunchKernelGGL(mc::quantize_kernel<4>, dim3(gx, njobs), dim3(256), 0, stream, dj, njobs, rounding);
    return mc::check_launch();
}

mc_status mc_attn_absmax(const void *x_fp16, long long count, unsigned *out_bits, hipStream_t stream) {
    if (count < 0 || !out_bits) return MC_EINVAL_ARG;
    if (count == 0) return MC_OK;
    if (!x_fp16) return MC_EINVAL_ARG;
    hipLaunchKernelGGL(mc::absmax_kernel, dim3(grid_for(count, 256 * 16)), dim3(256), 0, stream,
                       static_cast<const __half *>(x_fp16), count, out_bits);
    return mc::check_launch();
}

mc_status mc_attn_dequant(const int *in, int rows, int cols, float scale, const float *row_scale, float *out,
                          hipStream_t stream) {
    if (rows < 0 || cols <= 0) return MC_EINVAL_SHAPE;
    const long long count = (long long)rows * cols;
    if (count == 0) return MC_OK;
    if (!in || !out) return MC_EINVAL_ARG;
    hipLaunchKernelGGL(mc::dequant_kernel, dim3(grid_for(count, 256 * 8)), dim3(256), 0, stream, in, count, cols,
                       scale, row_scale, out);
    return mc::check_launch();
}

mc_status mc_attn_dequant_f16(const int *in, long long count, float divisor, void *out_f16, hipStream_t stream) {
    if (count < 0) return MC_EINVAL_SHAPE;
    if (count == 0) return MC_OK;
    if (!in || !out_f16) return MC_EINVAL_ARG;
    hipLaunchKernelGGL(mc::dequant_f16_kernel, dim3(grid_for(count, 256 * 8)), dim3(256), 0, stream, in, count,
                       divisor, static_cast<__half *>(out_f16));
    return mc::check_launch();
}

mc_status mc_attn_softmax_f16(int batch, int rows, int vec_length, const int *row_indices, const int *row_offsets,
                              const void *values_f16, long long values_stride, float sqrt_dk, float scale, int bits,
                              void *attn, long long attn_stride_bytes, hipStream_t stream) {
    if (bits != 16 && bits != 8 && bits != 4) return MC_EUNSUPPORTED;
    if (!mc::valid_v(vec_length)) return MC_EINVAL_V;
    if (rows < 0 || batch < 1 || batch > 65535 || values_stride < 0 || attn_stride_bytes < 0) return MC_EINVAL_SHAPE;
    if (rows == 0) return MC_OK;
    if (!row_offsets || !values_f16 || !attn) return MC_EINVAL_ARG;
    // 16-byte loads of the eight fp16 logits of a slot
    if (vec_length == 8 && ((reinterpret_cast<uintptr_t>(values_f16) & 15u) != 0 || values_stride % 8 != 0))
        return MC_EINVAL_ARG;
    constexpr int WPB = 4;
    const dim3 grid((unsigned)((rows + WPB - 1) / WPB), (unsigned)batch), block(64 * WPB);
    const auto *vals = static_cast<const __half *>(values_f16);
    auto *out = static_cast<unsigned char *>(attn);
#define MC_REF_SOFTMAX(V_, B_)                                                                                     \
    hipLaunchKernelGGL((mc::ref_softmax_kernel<V_, B_, WPB>), grid, block, 0, stream, rows, row_indices,           \
                       row_offsets, vals, values_stride, sqrt_dk, scale, out, attn_stride_bytes)
#define MC_REF_SOFTMAX_V(B_)                                                                                       \
    switch (vec_length) {                                                                                          \
    case 2: MC_REF_SOFTMAX(2, B_); break;                                                                          \
    case 4: MC_REF_SOFTMAX(4, B_); break;                                                                          \
    default: MC_REF_SOFTMAX(8, B_); break;                                                                         \
    }
    if (bits == 16) { MC_REF_SOFTMAX_V(16) }
    else if (bits == 8) { MC_REF_SOFTMAX_V(8) }
    else { MC_REF_SOFTMAX_V(4) }
#undef MC_REF_SOFTMAX_V
#undef MC_REF_SOFTMAX
    return mc::check_launch();
}

mc_status mc_attn_softmax_requant(int m_vec, int vec_length, const int *row_indices, const int *row_offsets,
                                  const int *s_vals, float alpha, float p_scale, int p_mode, void *p_out,
                                  float *row_inv_l, hipStream_t stream) {
    if (p_mode != MC_P_REFERENCE && p_mode != MC_P_ROW_SCALED && p_mode != MC_P_REFERENCE_4BIT) return MC_EINVAL_ARG;
    if (!mc::valid_v(vec_length)) return MC_EINVAL_V;
    if (m_vec < 0) return MC_EINVAL_SHAPE;
    if (m_vec == 0) return MC_OK;
    if (!row_offsets || !s_vals || !p_out) return MC_EINVAL_ARG;
    constexpr int WPB = 4;
    const dim3 grid((m_vec + WPB - 1) / WPB), block(64 * WPB);
    auto *po = static_cast<unsigned char *>(p_out);
    switch (vec_length) {
    case 2: hipLaunchKernelGGL((mc::softmax_requant_kernel<2, WPB>), grid, block, 0, stream, m_vec, row_indices, row_offsets, s_vals, alpha, p_scale, p_mode, po, row_inv_l); break;
    case 4: hipLaunchKernelGGL((mc::softmax_requant_kernel<4, WPB>), grid, block, 0, stream, m_vec, row_indices, row_offsets, s_vals, alpha, p_scale, p_mode, po, row_inv_l); break;
    default: hipLaunchKernelGGL((mc::softmax_requant_kernel<8, WPB>), grid, block, 0, stream, m_vec, row_indices, row_offsets, s_vals, alpha, p_scale, p_mode, po, row_inv_l); break;
    }
    return mc::check_launch();
}

mc_status mc_attn_forward(const mc_attn_matrix *matrices, const int *items, int n_items, int qk_bits, int v_bits,
                          int vec_length, float p_scale, int p_mode, int grid_waves, int debug, hipStream_t stream) {
    if (qk_bits != 8 && qk_bits != 4) return MC_EUNSUPPORTED;
    if (v_bits != 8 && !(v_bits == 4 && qk_bits == 4)) return MC_EUNSUPPORTED;
    if (p_mode != MC_P_REFERENCE && p_mode != MC_P_ROW_SCALED && p_mode != MC_P_REFERENCE_4BIT) return MC_EINVAL_ARG;
    if (!mc::valid_v(vec_length)) return MC_EINVAL_V;
    if (n_items < 0 || grid_waves < 0) return MC_EINVAL_ARG;
    if (n_items == 0) return MC_OK;
    if (!matrices || !items) return MC_EINVAL_ARG;
    constexpr int WPB = 4;
    long long waves = grid_waves > 0 ? grid_waves : n_items;
    if (waves > n_items) waves = n_items;
    const dim3 grid((unsigned)((waves + WPB - 1) / WPB)), block(64 * WPB);
    const auto *mats = reinterpret_cast<const mc::AttnMatrix *>(matrices);
    const auto *its = reinterpret_cast<const int2 *>(items);
#define MC_ATTN_LAUNCH(V_, B_, VB_, PM_)                                                                             \
    do {                                                                                                             \
        if (debug)                                                                                                   \
            hipLaunchKernelGGL((mc::attention_kernel<V_, B_, VB_, PM_, WPB, true>), grid, block, 0, stream, mats,    \
                               its, n_items, p_scale, p_mode);                                                       \
        else                                                                                                         \
            hipLaunchKernelGGL((mc::attention_kernel<V_, B_, VB_, PM_, WPB, false>), grid, block, 0, stream, mats,   \
                               its, n_items, p_scale, p_mode);                                                       \
    } while (0)
#define MC_ATTN_LAUNCH_P(V_, B_, VB_)                                                                                \
    switch (p_mode) {                                                                                                \
    case MC_P_REFERENCE: MC_ATTN_LAUNCH(V_, B_, VB_, 0); break;                                                      \
    case MC_P_ROW_SCALED: MC_ATTN_LAUNCH(V_, B_, VB_, 1); break;                                                     \
    default: MC_ATTN_LAUNCH(V_, B_, VB_, 2); break;                                                                  \
    }
#define MC_ATTN_LAUNCH_V(B_, VB_)                                                                                    \
    switch (vec_length) {                                                                                            \
    case 2: MC_ATTN_LAUNCH_P(2, B_, VB_); break;                                                                     \
    case 4: MC_ATTN_LAUNCH_P(4, B_, VB_); break;                                                                     \
    default: MC_ATTN_LAUNCH_P(8, B_, VB_); break;                                                                    \
    }
    if (qk_bits == 8) { MC_ATTN_LAUNCH_V(8, 8) }
    else if (v_bits == 8) { MC_ATTN_LAUNCH_V(4, 8) }
    else { MC_ATTN_LAUNCH_V(4, 4) }
#undef MC_ATTN_LAUNCH_P
#undef MC_ATTN_LAUNCH_V
#undef MC_ATTN_LAUNCH
    return mc::check_launch();
}

}  // extern "C"
