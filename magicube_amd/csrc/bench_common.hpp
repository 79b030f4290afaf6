// Shared pieces of the reference-compatible benchmark drivers
// (spmm_benchmark, sddmm_benchmark): sparse-pattern input, the reference's
// value generation, device buffers and event timing.
#pragma once

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/magicube_hip.h"

namespace mcb {

// Plain vector-sparse pattern: m_vec rows of V x 1 column vectors over k columns.
struct Pattern {
    int m_vec = 0;
    int k = 0;
    std::vector<int> row_offsets;   // m_vec + 1
    std::vector<int> col_indices;   // nnz_vec
    int nnz() const { return row_offsets.empty() ? 0 : row_offsets.back(); }
};

// DLMC .smtx: "m_vec, k, nnz_vec" then m_vec + 1 row offsets, then nnz_vec
// column indices, whitespace separated (the reader of
// SpMM/SpMM/spmm_benchmark.cpp:125-175).
inline Pattern read_smtx(const std::string &path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open " + path);
    std::string header;
    std::getline(in, header);
    for (char &c : header)
        if (c == ',') c = ' ';
    std::istringstream hs(header);
    Pattern p;
    long long nnz = 0;
    if (!(hs >> p.m_vec >> p.k >> nnz)) throw std::runtime_error("bad .smtx header in " + path);
    p.row_offsets.resize(p.m_vec + 1);
    for (int &x : p.row_offsets)
        if (!(in >> x)) throw std::runtime_error("short row offsets in " + path);
    p.col_indices.resize(nnz);
    for (int &x : p.col_indices)
        if (!(in >> x)) throw std::runtime_error("short column indices in " + path);
    if (p.row_offsets.front() != 0 || p.row_offsets.back() != nnz) throw std::runtime_error("inconsistent " + path);
    return p;
}

// "synthetic:M,K,sparsity,seed": nonzero vectors per V-row block ~
// Poisson((1 - sparsity) * K) clamped to [0, K], distinct columns uniform,
// sorted.  Used where the DLMC files are not available.
inline bool parse_synthetic(const std::string &spec, int V, Pattern &p) {
    const std::string tag = "synthetic:";
    if (spec.compare(0, tag.size(), tag) != 0) return false;
    std::string rest = spec.substr(tag.size());
    for (char &c : rest)
        if (c == ',') c = ' ';
    std::istringstream ss(rest);
    long long M = 0, K = 0, seed = 1;
    double sparsity = 0.9;
    if (!(ss >> M >> K >> sparsity)) throw std::runtime_error("synthetic:M,K,sparsity[,seed]");
    ss >> seed;
    if (M <= 0 || K <= 0 || M % V != 0 || sparsity < 0.0 || sparsity > 1.0)
        throw std::runtime_error("bad synthetic spec " + spec);
    std::mt19937_64 rng((uint64_t)seed);
    std::poisson_distribution<long long> count((1.0 - sparsity) * (double)K);
    p.m_vec = (int)(M / V);
    p.k = (int)K;
    p.row_offsets.assign(p.m_vec + 1, 0);
    p.col_indices.clear();
    std::vector<int> perm(K);
    for (int i = 0; i < p.m_vec; ++i) {
        const long long c = std::min<long long>(count(rng), K);
        // partial Fisher-Yates: the first c entries are a uniform c-subset
        for (long long j = 0; j < K; ++j) perm[j] = (int)j;
        for (long long j = 0; j < c; ++j) {
            std::uniform_int_distribution<long long> pick(j, K - 1);
            std::swap(perm[j], perm[pick(rng)]);
        }
        std::sort(perm.begin(), perm.begin() + c);
        p.col_indices.insert(p.col_indices.end(), perm.begin(), perm.begin() + c);
        p.row_offsets[i + 1] = (int)p.col_indices.size();
    }
    return true;
}

inline Pattern load_pattern(const std::string &spec, int V) {
    Pattern p;
    if (parse_synthetic(spec, V, p)) return p;
    return read_smtx(spec);
}

// The reference benchmarks' operand values (MakeDenseMatrix,
// SpMM/SpMM/include/bm_test_utils.h:130-164): a default-constructed
// std::default_random_engine copied into each fill, draws of
// uniform_real_distribution<float>(0, 1) mapped per word width —
// 1 byte: 127*u, 2 bytes: 10240*u, 4 bytes: 0x11111111 + int(117*u),
// 8 bytes: 0x1111111111111111 + int(117*u); the 4- and 8-byte words consume
// one extra draw first.  Words are little-endian, lane 0 in the low bits.
inline void fill_words(std::default_random_engine gen, int word_bytes, size_t words, void *out) {
    std::uniform_real_distribution<float> u(0.0f, 1.0f);
    for (size_t i = 0; i < words; ++i) {
        switch (word_bytes) {
        case 1: {
            const float t = 127.0 * u(gen);
            static_cast<signed char *>(out)[i] = (signed char)t;
            break;
        }
        case 2: {
            const float t = 10240.0 * u(gen);
            static_cast<short *>(out)[i] = (short)t;
            break;
        }
        case 4: {
            (void)u(gen);
            static_cast<int *>(out)[i] = 0x11111111 + int(117.0 * u(gen));
            break;
        }
        default: {
            (void)u(gen);
            static_cast<long long *>(out)[i] = 0x1111111111111111LL + int(117.0 * u(gen));
            break;
        }
        }
    }
}

// Lane `lane` of a packed little-endian array of `bits`-wide unsigned lanes.
inline uint32_t lane_at(const uint8_t *packed, size_t lane, int bits) {
    const size_t bit = lane * (size_t)bits;
    const size_t byte = bit >> 3;
    if (bits == 16) return (uint32_t)packed[byte] | (uint32_t)packed[byte + 1] << 8;
    if (bits == 8) return packed[byte];
    return (packed[byte] >> (bit & 7)) & 0xF;
}

inline void check_hip(hipError_t e, const char *what) {
    if (e != hipSuccess) {
        std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
        std::exit(1);
    }
}

inline void check_mc(mc_status s, const char *what) {
    if (s != MC_OK) {
        std::fprintf(stderr, "%s: %s\n", what, mc_status_string(s));
        std::exit(1);
    }
}

template <typename T> struct DeviceBuffer {
    T *ptr = nullptr;
    explicit DeviceBuffer(size_t count) { check_hip(hipMalloc(&ptr, std::max<size_t>(count, 1) * sizeof(T)), "hipMalloc"); }
    DeviceBuffer(const T *host, size_t count) : DeviceBuffer(count) {
        if (count) check_hip(hipMemcpy(ptr, host, count * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
    }
    ~DeviceBuffer() { (void)hipFree(ptr); }
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
};

// Mean per-launch time in ms: `warmup` untimed launches, then `runs`
// launches each bracketed by its own events, as the reference drivers time.
template <typename F> float time_launches(F launch, int warmup, int runs) {
    for (int i = 0; i < warmup; ++i) launch();
    check_hip(hipDeviceSynchronize(), "warm-up");
    hipEvent_t a, b;
    check_hip(hipEventCreate(&a), "hipEventCreate");
    check_hip(hipEventCreate(&b), "hipEventCreate");
    double total = 0.0;
    for (int i = 0; i < runs; ++i) {
        float ms = 0.0f;
        check_hip(hipEventRecord(a, nullptr), "hipEventRecord");
        launch();
        check_hip(hipEventRecord(b, nullptr), "hipEventRecord");
        check_hip(hipEventSynchronize(b), "hipEventSynchronize");
        check_hip(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
        total += ms;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return (float)(total / runs);
}

}  // namespace mcb
