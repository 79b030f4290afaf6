// sddmm_benchmark: the reference SDDMM benchmark's command line and report
// lines (SDDMM/SDDMM/sddmm_benchmark.cpp:320-351 for argv; :84, :116, :252-255
// and :275-277 for the lines its plot scripts parse), driving the MI355X
// library.
//
//   sddmm_benchmark bm k v sort function sparse preA preB
//
// bm is a DLMC .smtx file (its column count is the SDDMM's n) or
// "synthetic:M,N,sparsity[,seed]"; k in {64, 128, 256}; (preA, preB) in
// {(16,16), (8,8), (4,4)}.  Rows are padded to 8 slots and the output uses the
// reference layout [(j/8)][v][j%8]; `function` 1 verifies every output slot
// (padding slots must be 0) against a host loop.
#include <iostream>

#include "bench_common.hpp"

namespace {

using mcb::Pattern;

int run(const std::string &bm, int K, int V, int sorted, int func, int sparse, int preA, int preB) {
    if (V != 2 && V != 4 && V != 8) {
        std::printf("Unsupported vec_length!\n");
        return 1;
    }
    if (!((preA == 16 && preB == 16) || (preA == 8 && preB == 8) || (preA == 4 && preB == 4)) || sparse != 1) {
        std::printf("Unsupported precision!\n");
        return 1;
    }
    const int bits = preA;
    const Pattern pat = mcb::load_pattern(bm, V);
    const int m_vec = pat.m_vec, M = m_vec * V, N = pat.k, nnz = pat.nnz();
    std::printf("PreA: %d, PreB: %d, vec_len: %d, M: %d, M_vec: %d, N: %d, nnz: %d, K: %d\n", preA, preB, V, M, m_vec,
                N, nnz * V, K);
    std::vector<int> pairs(2 * (size_t)std::max(m_vec, 1));
    const int aligned = mc_format_align_rows(m_vec, pat.row_offsets.data(), 8, pairs.data());
    if (aligned < 0) {
        std::fprintf(stderr, "invalid row offsets\n");
        return 1;
    }
    std::cout << " nonzero_vec: " << nnz << " aligned_ nonzero_vec: " << aligned << "\n";
    std::vector<int> cols(std::max(aligned, 1));
    mcb::check_mc(mc_format_align_cols(m_vec, pat.row_offsets.data(), pat.col_indices.data(), pairs.data(), aligned,
                                       0, cols.data()),
                  "mc_format_align_cols");
    const size_t lhs_bytes = (size_t)M * K * bits / 8, rhs_bytes = (size_t)N * K * bits / 8;
    std::vector<uint8_t> lhs(lhs_bytes + 8), rhs(rhs_bytes + 8);
    std::default_random_engine gen;
    mcb::fill_words(gen, 4, lhs_bytes / 4, lhs.data());
    mcb::fill_words(gen, 4, rhs_bytes / 4, rhs.data());
    std::vector<int> order(std::max(m_vec, 1));
    if (sorted) mcb::check_mc(mc_format_sorted_rows(m_vec, pat.row_offsets.data(), order.data()), "sorted rows");
    else
        for (int i = 0; i < m_vec; ++i) order[i] = i;

    mcb::DeviceBuffer<int> d_pairs(pairs.data(), 2 * (size_t)m_vec), d_cols(cols.data(), aligned),
        d_order(order.data(), m_vec), d_out((size_t)aligned * V);
    mcb::DeviceBuffer<uint8_t> d_lhs(lhs.data(), lhs_bytes), d_rhs(rhs.data(), rhs_bytes);
    auto fn = bits == 16 ? mc_sddmm_l16r16 : bits == 8 ? mc_sddmm_l8r8 : mc_sddmm_l4r4;
    auto launch = [&]() {
        mcb::check_mc(fn(m_vec, K, N, d_order.ptr, d_pairs.ptr, d_cols.ptr, reinterpret_cast<const int *>(d_lhs.ptr),
                         reinterpret_cast<const int *>(d_rhs.ptr), d_out.ptr, V, MC_UNSIGNED, nullptr),
                      "sddmm");
    };
    if (bits == 4) std::printf("Using WMMA \n");
    const float ms = mcb::time_launches(launch, 32, 512);
    const double flops = 2.0 * nnz * V * K;
    std::cout << "Magicube SDDMM runtime " << ms << " ms" << "\n";
    if (func)
        std::cout << "SDDMM TOPS: " << flops / 1e12 << "  performance TOP/s: " << flops / 1e12 / (ms / 1000.0) << "\n";

    if (func) {
        std::vector<uint32_t> got((size_t)aligned * V);
        mcb::check_hip(hipMemcpy(got.data(), d_out.ptr, got.size() * 4, hipMemcpyDeviceToHost), "copy out");
        std::vector<uint32_t> exp(got.size(), 0u);
        for (int i = 0; i < m_vec; ++i)
            for (int j = pairs[2 * i]; j < pairs[2 * i + 1]; ++j) {
                const size_t col = (size_t)cols[j];
                for (int v = 0; v < V; ++v) {
                    uint32_t s = 0;
                    const size_t row = (size_t)i * V + v;
                    for (int kk = 0; kk < K; ++kk)
                        s += mcb::lane_at(lhs.data(), row * K + kk, bits) * mcb::lane_at(rhs.data(), col * K + kk, bits);
                    exp[(size_t)(j / 8) * 8 * V + (size_t)v * 8 + j % 8] = s;
                }
            }
        long long errors = 0;
        for (size_t i = 0; i < got.size(); ++i) errors += got[i] != exp[i];
        if (errors > 0)
            std::printf("SDDMM does not agree with SEQUENTIAL! Total %d, %lld errors!\n", aligned * V, errors);
        else std::printf("SDDMM results verification: PASS\n");
    }
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 9) {
        std::printf("usage: ./sddmm_benchmark [bm] [k] [v] [sort] [function] [sparse] [preA] [preB]\n");
        std::printf("bm      : DLMC .smtx file, or synthetic:M,N,sparsity[,seed]\n");
        std::printf("k       : reduction length, 64, 128 or 256\n");
        std::printf("v       : vector length, 2, 4 or 8\n");
        std::printf("sort    : 1 = longest rows first, 0 = in order\n");
        std::printf("function: 1 = verify every output, 0 = skip\n");
        std::printf("sparse  : 1 (SDDMM)\n");
        std::printf("preA preB: 16 16 | 8 8 | 4 4\n");
        return argc == 2 ? 0 : 1;
    }
    const std::string bm(argv[1]);
    std::cout << "Sparse matrix: " << bm << "\n";
    int rc = 0;
    try {
        rc = run(bm, std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5]),
                 std::atoi(argv[6]), std::atoi(argv[7]), std::atoi(argv[8]));
    } catch (const std::exception &e) {
        std::fprintf(stderr, "%s\n", e.what());
        rc = 1;
    }
    std::printf("\n");
    return rc;
}
