// Device building blocks for the gfx950 (CDNA4) vector-sparse kernels.
//
// Wave64 MFMA operand conventions used throughout (v_mfma_i32_16x16x64_i8):
//   lane l = r + 16*g   (r = l % 16, g = l / 16)
//   A operand: 16 int8 of row r, reduction slots [16g, 16g+16)
//   B operand: 16 int8 of column r, the same reduction slots
//   C/D:       4 int32 per lane, rows 4g+i (i = 0..3) of column r
// Only the pairing "A byte b of lane group g <-> B byte b of lane group g"
// matters for the sum, so every loader below places the byte for reduction
// slot s of its group at the same byte position in A and in B.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mc {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v2i __attribute__((ext_vector_type(2)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef unsigned int u32;

__device__ __forceinline__ v4i mfma_i8(v4i a, v4i b, v4i c) {
    return __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ u32 perm(u32 hi, u32 lo, u32 sel) {
    // byte selector over the 8-byte value {hi:lo}; 0..3 pick lo, 4..7 pick hi
    return __builtin_amdgcn_perm(hi, lo, sel);
}

// 4x4 byte transpose: r[k] holds 4 bytes (columns 0..3) of row k; o[c] gets
// column c as bytes (row0, row1, row2, row3).
__device__ __forceinline__ void transpose4x4(u32 r0, u32 r1, u32 r2, u32 r3, u32 o[4]) {
    const u32 t01l = perm(r1, r0, 0x05010400u);   // r0b0 r1b0 r0b1 r1b1
    const u32 t01h = perm(r1, r0, 0x07030602u);   // r0b2 r1b2 r0b3 r1b3
    const u32 t23l = perm(r3, r2, 0x05010400u);
    const u32 t23h = perm(r3, r2, 0x07030602u);
    o[0] = perm(t23l, t01l, 0x05040100u);
    o[1] = perm(t23l, t01l, 0x07060302u);
    o[2] = perm(t23h, t01h, 0x05040100u);
    o[3] = perm(t23h, t01h, 0x07060302u);
}

// Nibble helpers on packed dwords (8 nibbles, nibble 0 in bits 0..3).
__device__ __forceinline__ u32 nib_even(u32 x) { return x & 0x0F0F0F0Fu; }          // nibbles 0,2,4,6
__device__ __forceinline__ u32 nib_odd(u32 x) { return (x >> 4) & 0x0F0F0F0Fu; }    // nibbles 1,3,5,7
// Sign-extend four 4-bit values held in the low nibble of each byte.
__device__ __forceinline__ u32 sext4_bytes(u32 t) {
    const u32 m = t & 0x08080808u;
    return t | ((m << 5) - (m << 1));   // 30*m == 0xF0 per flagged byte, no carries
}

// 4 x 4 transpose of ints within each quad of lanes: on return lane q
// (q = lane & 3) holds in x[k] what lane k of its quad held in x[q].  Two
// butterfly stages (lane bit 0 with register bit 0, then bit 1 with bit 1),
// each one DPP quad_perm swap per register pair.
__device__ __forceinline__ void quad_transpose4(int (&x)[4], int lane) {
    // written out with selects: a data-dependent choice of array element
    // would move the array out of registers
    const bool h0 = (lane & 1) != 0, h1 = (lane & 2) != 0;
    const int g0 = __builtin_amdgcn_update_dpp(0, h0 ? x[0] : x[1], 0xB1, 0xF, 0xF, false);   // quad_perm [1,0,3,2]
    const int g1 = __builtin_amdgcn_update_dpp(0, h0 ? x[2] : x[3], 0xB1, 0xF, 0xF, false);
    const int y0 = h0 ? g0 : x[0], y1 = h0 ? x[1] : g0, y2 = h0 ? g1 : x[2], y3 = h0 ? x[3] : g1;
    const int k0 = __builtin_amdgcn_update_dpp(0, h1 ? y0 : y2, 0x4E, 0xF, 0xF, false);      // quad_perm [2,3,0,1]
    const int k1 = __builtin_amdgcn_update_dpp(0, h1 ? y1 : y3, 0x4E, 0xF, 0xF, false);
    x[0] = h1 ? k0 : y0;
    x[1] = h1 ? k1 : y1;
    x[2] = h1 ? y2 : k0;
    x[3] = h1 ? y3 : k1;
}

__device__ __forceinline__ int sum_bytes(v4i a) {
    int s = __builtin_amdgcn_sdot4(a[0], 0x01010101, 0, false);
    s = __builtin_amdgcn_sdot4(a[1], 0x01010101, s, false);
    s = __builtin_amdgcn_sdot4(a[2], 0x01010101, s, false);
    s = __builtin_amdgcn_sdot4(a[3], 0x01010101, s, false);
    return s;
}

// Raw buffer resource over [base, base + bytes): loads whose offset falls
// outside return zero and touch no memory (gfx9 descriptor word 3).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void *base, u32 bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ u32 buf_load_b32(__amdgpu_buffer_rsrc_t r, u32 off) {
    return (u32)__builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0);
}
__device__ __forceinline__ u32 buf_load_u16(__amdgpu_buffer_rsrc_t r, u32 off) {
    return (u32)__builtin_amdgcn_raw_buffer_load_b16(r, off, 0, 0);
}
__device__ __forceinline__ v2i buf_load_b64(__amdgpu_buffer_rsrc_t r, u32 off) {
    return __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
}
__device__ __forceinline__ v4i buf_load_b128(__amdgpu_buffer_rsrc_t r, u32 off) {
    return __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
}

// ds_read_b64_tr_b8: within each 16-lane group, lane 8h + j
// receives byte j of the 8-byte segments addressed by lanes h, h + 2, ...,
// h + 14 (8 bytes: one column of an 8 x 8 byte block).
typedef __attribute__((address_space(3))) v2i lds_v2i;
__device__ __forceinline__ v2i lds_read_tr8(const unsigned char *base, u32 off) {
    return __builtin_amdgcn_ds_read_tr8_b64_v2i32((lds_v2i *)(base + off));
}

// Wave-uniform copies of values every lane loaded from the same address.
// Values in VGPRs cannot feed a buffer resource or a scalar branch without a
// per-lane waterfall loop; one v_readfirstlane each makes them SGPR-resident.
__device__ __forceinline__ int uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ float uniform(float x) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}
template <typename T> __device__ __forceinline__ T *uniform(T *p) {
    const unsigned long long v = reinterpret_cast<unsigned long long>(p);
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)v);
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(v >> 32));
    return reinterpret_cast<T *>(((unsigned long long)hi << 32) | lo);
}

// Cross-lane read of lane (self + delta) within the wave (delta multiple of 16).
__device__ __forceinline__ int lane_down(int x, int delta) {
    const int lane = __lane_id();
    return __builtin_amdgcn_ds_bpermute(((lane + delta) & 63) << 2, x);
}
// Lane p of each row of 16 lanes, broadcast to that row (DPP row_newbcast,
// no LDS traffic); p must fold to a constant.
__device__ __forceinline__ int row_bcast16(int x, int p) {
    switch (p & 15) {
    case 0: return __builtin_amdgcn_update_dpp(0, x, 0x150, 0xF, 0xF, false);
    case 1: return __builtin_amdgcn_update_dpp(0, x, 0x151, 0xF, 0xF, false);
    case 2: return __builtin_amdgcn_update_dpp(0, x, 0x152, 0xF, 0xF, false);
    case 3: return __builtin_amdgcn_update_dpp(0, x, 0x153, 0xF, 0xF, false);
    case 4: return __builtin_amdgcn_update_dpp(0, x, 0x154, 0xF, 0xF, false);
    case 5: return __builtin_amdgcn_update_dpp(0, x, 0x155, 0xF, 0xF, false);
    case 6: return __builtin_amdgcn_update_dpp(0, x, 0x156, 0xF, 0xF, false);
    case 7: return __builtin_amdgcn_update_dpp(0, x, 0x157, 0xF, 0xF, false);
    case 8: return __builtin_amdgcn_update_dpp(0, x, 0x158, 0xF, 0xF, false);
    case 9: return __builtin_amdgcn_update_dpp(0, x, 0x159, 0xF, 0xF, false);
    case 10: return __builtin_amdgcn_update_dpp(0, x, 0x15A, 0xF, 0xF, false);
    case 11: return __builtin_amdgcn_update_dpp(0, x, 0x15B, 0xF, 0xF, false);
    case 12: return __builtin_amdgcn_update_dpp(0, x, 0x15C, 0xF, 0xF, false);
    case 13: return __builtin_amdgcn_update_dpp(0, x, 0x15D, 0xF, 0xF, false);
    case 14: return __builtin_amdgcn_update_dpp(0, x, 0x15E, 0xF, 0xF, false);
    default: return __builtin_amdgcn_update_dpp(0, x, 0x15F, 0xF, 0xF, false);
    }
}
__device__ __forceinline__ int lane_read(int x, int src_lane) {
    return __builtin_amdgcn_ds_bpermute((src_lane & 63) << 2, x);
}

}  // namespace mc
