// Memory-access floors of the headline SpMM's dataflows on one GPU
// (M = K = 4096, N = 512, V = 8, Poisson nonzero count per vector row around
// 409.6, int8 B): the same loads as the kernels, with no arithmetic, so a
// kernel trace of this program shows what the loads alone cost.
//
//   gather         the gather kernel's B-row loads: one workgroup of four
//                  waves per (vector row, 256-column half), each lane one
//                  dword of the 16 B rows of its lane group's slots per
//                  64-slot k-block, gather offsets staged in LDS per 1024
//                  slots (spmm_kernel.hpp, tile config 17)
//   gather_8w      one workgroup of eight waves per vector row (all 512
//                  columns), otherwise the same (tile config 13, the default
//                  tile of this shape)
//   gather_2deep   the four-wave gather with two k-blocks of loads in flight
//   gather_x4      one workgroup per vector row, 16 bytes per lane
//   fill_strided   a B-stationary kernel's fill: 256 workgroups, each copying
//                  a 32-column slice of all K rows of row-major B (32 of
//                  every 512 bytes) into 128 KiB of LDS
//   fill_contig    the same slice stored contiguously (B relaid out)
//   chain_1wg      one workgroup doing one fill: the latency of a dependent
//                  load -> LDS -> barrier -> store chain
//
// Usage: gather_floor [reps]   (prints one JSON line per probe: the host-side
// mean per launch over reps back-to-back launches; kernel durations come
// from rocprofv3 --kernel-trace)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "mc_device.hpp"

using namespace mc;

#define GF_CHECK(x)                                                                       \
    do {                                                                                  \
        hipError_t e_ = (x);                                                              \
        if (e_ != hipSuccess) {                                                           \
            fprintf(stderr, "gather_floor: %s at line %d\n", hipGetErrorString(e_), __LINE__); \
            exit(1);                                                                      \
        }                                                                                 \
    } while (0)

namespace {

constexpr int kK = 4096, kN = 512, kMVec = 512;
constexpr u32 kOob = 0x40000000u;     // added to an offset: outside the 2 MiB B buffer

// LW dwords per lane, DEPTH k-blocks in flight; NW waves per workgroup cover
// NW * 64 * LW bytes of each B row (HALVES workgroups per vector row).
template <int LW, int DEPTH, int NW, int HALVES>
__global__ __launch_bounds__(64 * NW) void gather(const int *pairs, const int *cols, const unsigned char *B,
                                                  unsigned *sink) {
    __shared__ u32 s_off[1024];
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int r = lane & 15, g = lane >> 4;
    const int row = blockIdx.x / HALVES, half = blockIdx.x % HALVES;
    const int start = pairs[2 * row], end = pairs[2 * row + 1];
    const int aend = start + ((end - start + 15) & ~15);
    const u32 coff = (u32)(half * NW + wave) * 64u * LW + (u32)r * 4u * LW;
    const __amdgpu_buffer_rsrc_t rb = make_rsrc(B, (u32)kK * kN);
    u32 acc = 0;
    for (int c0 = start; c0 < aend; c0 += 1024) {
        const int cn = min(aend - c0, 1024);
        const int cblk = (cn + 63) >> 6;
        __syncthreads();
        for (int i = threadIdx.x; i < cblk * 64; i += 64 * NW) {
            u32 o = kOob;
            if (i < cn) {
                const int idx = cols[c0 + i];
                if (idx >= 0) o = (u32)idx * (u32)kN;
            }
            s_off[i] = o;
        }
        __syncthreads();
        for (int kb = 0; kb < cblk; kb += DEPTH) {
            u32 v[DEPTH][16][LW];
#pragma unroll
            for (int d = 0; d < DEPTH; ++d) {
                const int lb = min(kb + d, cblk - 1);
                const u32 add = coff + (kb + d < cblk ? 0u : kOob);
                const v4i *so = reinterpret_cast<const v4i *>(s_off + lb * 64 + 16 * g);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const v4i o4 = so[q];
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const u32 off = (u32)o4[e] + add;
                        if constexpr (LW == 1) {
                            v[d][4 * q + e][0] = buf_load_b32(rb, off);
                        } else {
                            const v4i x = buf_load_b128(rb, off);
#pragma unroll
                            for (int w = 0; w < 4; ++w) v[d][4 * q + e][w] = (u32)x[w];
                        }
                    }
                }
            }
#pragma unroll
            for (int d = 0; d < DEPTH; ++d)
#pragma unroll
                for (int i = 0; i < 16; ++i)
#pragma unroll
                    for (int w = 0; w < LW; ++w) acc = acc * 31u + v[d][i][w];
        }
    }
    sink[blockIdx.x * (64 * NW) + threadIdx.x] = acc;
}

// 128 KiB slice (32 bytes of each of the K rows) into LDS, then one read back.
template <bool CONTIG>
__global__ __launch_bounds__(1024) void fill(const unsigned char *B, unsigned *sink) {
    extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
    const int slice = blockIdx.x & 15;
    const __amdgpu_buffer_rsrc_t rb = make_rsrc(B, (u32)kK * kN);
    v4i x[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        const int c = threadIdx.x + 1024 * u;            // 16-byte chunk of the slice
        const u32 off = CONTIG ? (u32)slice * 131072u + (u32)c * 16u
                               : (u32)(c >> 1) * (u32)kN + (u32)slice * 32u + (u32)(c & 1) * 16u;
        x[u] = buf_load_b128(rb, off);
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) *reinterpret_cast<v4i *>(lds + (threadIdx.x + 1024 * u) * 16) = x[u];
    __syncthreads();
    const v4i y = *reinterpret_cast<const v4i *>(lds + ((threadIdx.x * 16 + 4096) & 131071));
    sink[blockIdx.x * 1024 + threadIdx.x] = (unsigned)(y[0] ^ y[1] ^ y[2] ^ y[3]);
}

template <class F> double mean_us(F launch, int reps) {
    for (int i = 0; i < 200; ++i) launch();
    GF_CHECK(hipDeviceSynchronize());
    hipEvent_t a, b;
    GF_CHECK(hipEventCreate(&a));
    GF_CHECK(hipEventCreate(&b));
    GF_CHECK(hipEventRecord(a));
    for (int i = 0; i < reps; ++i) launch();
    GF_CHECK(hipEventRecord(b));
    GF_CHECK(hipEventSynchronize(b));
    float ms = 0.f;
    GF_CHECK(hipEventElapsedTime(&ms, a, b));
    GF_CHECK(hipGetLastError());
    return ms * 1000.0 / reps;
}

}  // namespace

int main(int argc, char **argv) {
    const int reps = argc > 1 ? atoi(argv[1]) : 2000;
    std::mt19937 rng(7);
    std::poisson_distribution<int> pois(409.6);
    std::vector<int> pairs(2 * kMVec), cols, all(kK);
    for (int i = 0; i < kMVec; ++i) {
        const int n = std::min(pois(rng), kK);
        for (int j = 0; j < kK; ++j) all[j] = j;
        std::shuffle(all.begin(), all.end(), rng);
        std::sort(all.begin(), all.begin() + n);
        pairs[2 * i] = (int)cols.size();
        pairs[2 * i + 1] = (int)cols.size() + n;
        cols.insert(cols.end(), all.begin(), all.begin() + n);
        while (cols.size() % 16) cols.push_back(-1);
    }
    std::vector<unsigned char> hB((size_t)kK * kN);
    for (auto &x : hB) x = (unsigned char)rng();
    int *dp = nullptr, *dc = nullptr;
    unsigned char *dB = nullptr;
    unsigned *sink = nullptr;
    GF_CHECK(hipMalloc(&dp, pairs.size() * 4));
    GF_CHECK(hipMalloc(&dc, cols.size() * 4));
    GF_CHECK(hipMalloc(&dB, hB.size()));
    GF_CHECK(hipMalloc(&sink, (size_t)kMVec * 512 * 4 + 256 * 1024 * 4));
    GF_CHECK(hipMemcpy(dp, pairs.data(), pairs.size() * 4, hipMemcpyHostToDevice));
    GF_CHECK(hipMemcpy(dc, cols.data(), cols.size() * 4, hipMemcpyHostToDevice));
    GF_CHECK(hipMemcpy(dB, hB.data(), hB.size(), hipMemcpyHostToDevice));
    GF_CHECK(hipFuncSetAttribute(reinterpret_cast<const void *>(fill<true>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, 131072));
    GF_CHECK(hipFuncSetAttribute(reinterpret_cast<const void *>(fill<false>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, 131072));
    size_t nnz = 0;
    for (int i = 0; i < kMVec; ++i) nnz += pairs[2 * i + 1] - pairs[2 * i];
    printf("{\"workload\": \"M=K=4096 N=512 V=8 int8 B, Poisson 409.6 per vector row\", \"nnz_vec\": %zu, "
           "\"b_row_segments_64B\": %zu}\n", nnz, nnz * (kN / 64));
    auto line = [](const char *name, double us) { printf("{\"probe\": \"%s\", \"host_us_per_launch\": %.3f}\n", name, us); };
    line("gather", mean_us([&] { hipLaunchKernelGGL((gather<1, 1, 4, 2>), dim3(kMVec * 2), dim3(256), 0, 0, dp, dc, dB, sink); }, reps));
    line("gather_8w", mean_us([&] { hipLaunchKernelGGL((gather<1, 1, 8, 1>), dim3(kMVec), dim3(512), 0, 0, dp, dc, dB, sink); }, reps));
    line("gather_2deep", mean_us([&] { hipLaunchKernelGGL((gather<1, 2, 4, 2>), dim3(kMVec * 2), dim3(256), 0, 0, dp, dc, dB, sink); }, reps));
    line("gather_x4", mean_us([&] { hipLaunchKernelGGL((gather<4, 1, 2, 1>), dim3(kMVec), dim3(128), 0, 0, dp, dc, dB, sink); }, reps));
    line("fill_strided", mean_us([&] { hipLaunchKernelGGL(fill<false>, dim3(256), dim3(1024), 131072, 0, dB, sink); }, reps));
    line("fill_contig", mean_us([&] { hipLaunchKernelGGL(fill<true>, dim3(256), dim3(1024), 131072, 0, dB, sink); }, reps));
    line("chain_1wg", mean_us([&] { hipLaunchKernelGGL(fill<true>, dim3(1), dim3(1024), 131072, 0, dB, sink); }, reps));
    GF_CHECK(hipDeviceSynchronize());
    GF_CHECK(hipFree(dp));
    GF_CHECK(hipFree(dc));
    GF_CHECK(hipFree(dB));
    GF_CHECK(hipFree(sink));
    return 0;
}
