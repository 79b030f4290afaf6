// spmm_benchmark: the reference SpMM benchmark's command line and report
// lines (SpMM/SpMM/spmm_benchmark.cpp:757-791 for argv; :145-146, :203, :387,
// :634-637 and :729-731 for the lines its plot scripts parse), driving the
// MI355X library.
//
//   spmm_benchmark bm n v kernel sort function sparse preA preB
//
// bm is a DLMC .smtx file or "synthetic:M,K,sparsity[,seed]".  kernel 0 and
// sparse 1 select the SpMM (the reference's other kernels are baselines that
// the library does not carry); pairs (preA, preB) in {(8,8), (12,8), (16,8),
// (16,16), (4,4), (8,4), (12,4), (16,4)}, 12-bit lanes in 16-bit containers
// as the reference holds them.  Operand values follow the reference's
// generator (MakeDenseMatrix of the reference's TypeA for the pair), B is K x n, the
// lanes are unsigned as in the reference's MMAs, and `function` 1 verifies
// every output element against a host loop over the unpadded pattern.
#include <cinttypes>
#include <iostream>

#include "bench_common.hpp"

namespace {

using mcb::Pattern;

struct PairInfo {
    int lhs_bits;
    int rhs_bits;
    mc_status (*fn)(int, int, int, int, const int *, const int *, const int *, const int *, const int *, int *,
                    mc_signedness, hipStream_t);
};

bool pair_for(int preA, int preB, PairInfo &p) {
    if (preA == 8 && preB == 8) p = {8, 8, mc_spmm_l8r8};
    else if (preA == 16 && preB == 8) p = {16, 8, mc_spmm_l16r8};
    else if (preA == 8 && preB == 4) p = {8, 4, mc_spmm_l8r4};
    else if (preA == 4 && preB == 4) p = {4, 4, mc_spmm_l4r4};
    else if (preA == 12 && preB == 8) p = {12, 8, mc_spmm_l12r8};
    else if (preA == 16 && preB == 16) p = {16, 16, mc_spmm_l16r16};
    else if (preA == 12 && preB == 4) p = {12, 4, mc_spmm_l12r4};
    else if (preA == 16 && preB == 4) p = {16, 4, mc_spmm_l16r4};
    else return false;
    return true;
}

// Exact unsigned C = A B modulo 2^32 over the unpadded pattern.
std::vector<uint32_t> host_spmm(const Pattern &pat, int V, int N, int abits, int bbits, const uint8_t *a,
                                const uint8_t *b) {
    std::vector<uint32_t> c((size_t)pat.m_vec * V * N, 0u);
    std::vector<uint32_t> brow(N);
    for (int i = 0; i < pat.m_vec; ++i)
        for (int j = pat.row_offsets[i]; j < pat.row_offsets[i + 1]; ++j) {
            const size_t col = (size_t)pat.col_indices[j];
            for (int n = 0; n < N; ++n) brow[n] = mcb::lane_at(b, col * N + n, bbits);
            for (int v = 0; v < V; ++v) {
                // 12-bit lanes: low 12 bits of a 16-bit container
                const uint32_t av = abits == 12 ? mcb::lane_at(a, (size_t)j * V + v, 16) & 0xFFFu
                                                : mcb::lane_at(a, (size_t)j * V + v, abits);
                uint32_t *crow = c.data() + ((size_t)i * V + v) * N;
                for (int n = 0; n < N; ++n) crow[n] += av * brow[n];
            }
        }
    return c;
}

int run(const std::string &bm, int N, int V, int kernel, int sorted, int func, int sparse, int preA, int preB) {
    PairInfo pi;
    if (!pair_for(preA, preB, pi) || (V != 2 && V != 4 && V != 8)) {
        std::printf("Unsupported precision and vec_length!\n");
        return 1;
    }
    if (kernel != 0 || sparse != 1) {
        std::printf("Unsupported Kernel \n");
        return 1;
    }
    const Pattern pat = mcb::load_pattern(bm, V);
    const int m_vec = pat.m_vec, K = pat.k, nnz = pat.nnz();
    const int mma_k = mc_format_row_alignment(pi.lhs_bits, pi.rhs_bits);
    std::printf("preA %d, preA_cut %d, preB %d, vec_length %d \n", preA == 12 ? 16 : preA, preA, preB, V);
    std::printf("m_vec %d, dimN %d, nonzeros_vec %d, dimk %d, mma_k_dim %d \n", m_vec, N, nnz, K, mma_k);

    std::vector<int> pairs(2 * (size_t)std::max(m_vec, 1));
    const int aligned = mc_format_align_rows(m_vec, pat.row_offsets.data(), mma_k, pairs.data());
    if (aligned < 0) {
        std::fprintf(stderr, "invalid row offsets\n");
        return 1;
    }
    std::cout << " nonzero_vec: " << nnz << " aligned_ nonzero_vec: " << aligned << "\n";
    std::vector<int> cols(std::max(aligned, 1));
    mcb::check_mc(mc_format_align_cols(m_vec, pat.row_offsets.data(), pat.col_indices.data(), pairs.data(), aligned,
                                       pi.rhs_bits == 4, cols.data()),
                  "mc_format_align_cols");

    // values: nnz vectors of V lanes; B: K x N lanes (reference generator,
    // TypeA short for L16-R16, else the widest word up to 8 bytes:
    // SpMM/SpMM/spmm_benchmark.cpp:805-831)
    const int vec_bytes = V * (pi.lhs_bits == 12 ? 16 : pi.lhs_bits) / 8;
    const int word = pi.rhs_bits == 16 ? 2 : std::min(8, vec_bytes);
    std::vector<uint8_t> values((size_t)nnz * vec_bytes + 8);
    std::default_random_engine gen;
    mcb::fill_words(gen, word, (size_t)nnz * vec_bytes / word, values.data());
    const size_t b_bytes = (size_t)K * N * pi.rhs_bits / 8;
    std::vector<uint8_t> rhs(b_bytes + 8);
    mcb::fill_words(gen, 4, b_bytes / 4, rhs.data());

    double flops = 2.0 * nnz * V * N;
    if (func) std::cout << "total Gflops: " << flops / 1e9 << "\n";

    std::vector<uint8_t> prepared((size_t)aligned * vec_bytes + 16);
    mcb::check_mc(mc_format_prepare_values(pi.lhs_bits, pi.rhs_bits, V, m_vec, pat.row_offsets.data(), pairs.data(),
                                           aligned, values.data(), prepared.data()),
                  "mc_format_prepare_values");
    std::vector<int> order(std::max(m_vec, 1));
    if (sorted) mcb::check_mc(mc_format_sorted_rows(m_vec, pat.row_offsets.data(), order.data()), "sorted rows");
    else
        for (int i = 0; i < m_vec; ++i) order[i] = i;

    mcb::DeviceBuffer<int> d_pairs(pairs.data(), 2 * (size_t)m_vec), d_cols(cols.data(), aligned),
        d_order(order.data(), m_vec), d_out((size_t)m_vec * V * N);
    mcb::DeviceBuffer<uint8_t> d_vals(prepared.data(), prepared.size()), d_rhs(rhs.data(), b_bytes);
    auto launch = [&]() {
        mcb::check_mc(pi.fn(m_vec, V, N, K, d_order.ptr, d_pairs.ptr, d_cols.ptr,
                            reinterpret_cast<const int *>(d_vals.ptr), reinterpret_cast<const int *>(d_rhs.ptr),
                            d_out.ptr, MC_UNSIGNED, nullptr),
                      "spmm");
    };
    const float ms = mcb::time_launches(launch, 16, 512);
    std::cout << "Magicube SpMM runtime " << ms << " ms" << "\n";
    if (func) std::cout << "performance TOP/s: " << flops / 1e9 / (ms / 1000.0) / 1000.0 << "\n";

    if (func) {
        std::vector<uint32_t> got((size_t)m_vec * V * N);
        mcb::check_hip(hipMemcpy(got.data(), d_out.ptr, got.size() * 4, hipMemcpyDeviceToHost), "copy C");
        const std::vector<uint32_t> exp = host_spmm(pat, V, N, pi.lhs_bits, pi.rhs_bits, values.data(), rhs.data());
        long long errors = 0;
        for (size_t i = 0; i < got.size(); ++i) errors += got[i] != exp[i];
        if (errors > 0) std::printf("SPMM does not agree with SEQUENTIAL! %lld errors!\n", errors);
        else std::printf("Results verification: PASS\n");
    }
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 10) {
        std::printf("usage: ./spmm_benchmark [bm] [n] [v] [kernel] [sort] [function] [sparse] [preA] [preB]\n");
        std::printf("bm      : DLMC .smtx file, or synthetic:M,K,sparsity[,seed]\n");
        std::printf("n       : columns of B and C\n");
        std::printf("v       : vector length, 2, 4 or 8\n");
        std::printf("kernel  : 0 (MFMA SpMM)\n");
        std::printf("sort    : 1 = longest rows first, 0 = in order\n");
        std::printf("function: 1 = verify every output, 0 = skip\n");
        std::printf("sparse  : 1 (SpMM)\n");
        std::printf("preA preB: 8 8 | 12 8 | 16 8 | 16 16 | 4 4 | 8 4 | 12 4 | 16 4\n");
        return argc == 2 ? 0 : 1;
    }
    const std::string bm(argv[1]);
    std::cout << "Sparse matrix: " << bm << "\n";
    int rc = 0;
    try {
        rc = run(bm, std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5]),
                 std::atoi(argv[6]), std::atoi(argv[7]), std::atoi(argv[8]), std::atoi(argv[9]));
    } catch (const std::exception &e) {
        std::fprintf(stderr, "%s\n", e.what());
        rc = 1;
    }
    std::printf("\n");
    return rc;
}
