// baseline_spmm_benchmark: the command line and report lines of the
// reference's vendor-library SpMM baselines (baselines/spmm_benchmark.cpp:
// 780-823 for argv, :268-602 Blocked-ELL, :603-770 dense GEMM), driving
// rocSPARSE and hipBLASLt through libmagicube_vendor.so.
//
//   baseline_spmm_benchmark bm k v kernel sort function sparse mixed
//
// bm is a DLMC .smtx file or "synthetic:M,K,sparsity[,seed]" (M rows of the
// vector matrix); k is the dense dimension N.  Modes (the reference's
// launch_spmm_*.py settings):
//   sparse -1          int8 dense GEMM (hipBLASLt, B stored transposed)
//   sparse -2          fp16 dense GEMM (hipBLASLt)
//   sparse  3          int8 blocked SpMM: V x V blocks, uniform per block row,
//                      at the matrix's element density (rocSPARSE BSR; the
//                      reference's cuSPARSE Blocked-ELL)
//   sparse  2          the same in fp16 (fp32 output)
//   sparse  1 kernel 3 the vector matrix as scalar CSR, fp16 (rocSPARSE CSR)
// The reference's own fp16 vectorSparse kernels (sparse 1, kernel 0/1) are
// not carried.  Times are the mean over 512 launches each bracketed by its
// own events after 16 untimed ones, as the reference measures.  Values are
// uniform: int8 in [0, 127) (the reference's float draws in (-1, 1) truncate
// to zero for int8), fp16 in (-1, 1).  function 1 checks the result against a
// host computation: exactly for int8, to 1e-2 relative for fp16.
#include <hip/hip_fp16.h>

#include <cmath>
#include <iostream>

#include "../../include/magicube_vendor.h"
#include "bench_common.hpp"

namespace {

using mcb::Pattern;

void check_v(mc_status s, const char *what) {
    if (s != MC_OK) {
        std::fprintf(stderr, "%s: %s (%s)\n", what, mc_status_string(s), mcv_last_error());
        std::exit(s == MC_EUNSUPPORTED ? 3 : 1);
    }
}

std::vector<int8_t> int8_values(std::default_random_engine &gen, size_t n) {
    std::vector<int8_t> v(n);
    mcb::fill_words(gen, 1, n, v.data());
    gen.discard(n);
    return v;
}

std::vector<__half> half_values(std::default_random_engine &gen, size_t n) {
    std::uniform_real_distribution<float> u(-1.0f, 1.0f);
    std::vector<__half> v(n);
    for (auto &x : v) x = __float2half(u(gen));
    return v;
}

// Uniform V x V blocks: int(density * kb) + 1 distinct block columns per
// block row, ascending (the structure of bm_test_utils.h:47-67).
void uniform_blocks(int mb, int kb, double density, std::default_random_engine &gen, std::vector<int> &row_ptr,
                    std::vector<int> &col) {
    const int per = std::min(kb, (int)(density * kb) + 1);
    row_ptr.resize(mb + 1);
    col.clear();
    std::vector<int> idx(kb);
    for (int i = 0; i < mb; ++i) {
        for (int j = 0; j < kb; ++j) idx[j] = j;
        std::shuffle(idx.begin(), idx.end(), gen);
        std::sort(idx.begin(), idx.begin() + per);
        row_ptr[i] = (int)col.size();
        col.insert(col.end(), idx.begin(), idx.begin() + per);
    }
    row_ptr[mb] = (int)col.size();
}

template <typename F> void report(const char *label, F launch) {
    const float ms = mcb::time_launches(launch, 16, 512);
    std::cout << label << " runtime " << ms << " ms\n";
}

void verdict(size_t errors, const char *who) {
    if (errors) std::printf("%s does not agree with SEQUENTIAL! %zu errors!\n", who, errors);
    else std::printf("Results verified: they agree.\n");
}

int run(const std::string &bm, int N, int V, int kernel, int func, int sparse, int mixed) {
    const Pattern pat = mcb::load_pattern(bm, V);
    const int m = pat.m_vec * V, K = pat.k;
    std::cout << "vec_len: " << V << " m_vec: " << pat.m_vec << " m: " << m << " n: " << N << " k: " << K << "\n";
    std::default_random_engine gen;
    hipStream_t stream = nullptr;
    mcv_plan *plan = nullptr;

    if (sparse == -1 || sparse == -2) {
        const bool i8 = sparse == -1;
        std::printf(i8 ? "8-bit dense Baseline\n" : "Dense Baseline\n");
        if (i8) {
            const auto a = int8_values(gen, (size_t)m * K), bt = int8_values(gen, (size_t)N * K);
            mcb::DeviceBuffer<int8_t> da(a.data(), a.size()), db(bt.data(), bt.size());
            mcb::DeviceBuffer<int32_t> dc((size_t)m * N);
            check_v(mcv_gemm_create(MCV_I8, m, K, N, da.ptr, db.ptr, 1, dc.ptr, stream, &plan), "hipBLASLt int8");
            report("hipBLASLt int8 dense for spmm", [&] { check_v(mcv_run(plan), "mcv_run"); });
            if (func) {
                std::vector<int32_t> c((size_t)m * N);
                mcb::check_hip(hipMemcpy(c.data(), dc.ptr, c.size() * 4, hipMemcpyDeviceToHost), "copy");
                size_t err = 0;
                for (int i = 0; i < m; ++i)
                    for (int j = 0; j < N; ++j) {
                        long long s = 0;
                        for (int t = 0; t < K; ++t) s += (long long)a[(size_t)i * K + t] * bt[(size_t)j * K + t];
                        err += (int32_t)s != c[(size_t)i * N + j];
                    }
                verdict(err, "hipBLASLt");
            }
        } else {
            const auto a = half_values(gen, (size_t)m * K), bt = half_values(gen, (size_t)N * K);
            mcb::DeviceBuffer<__half> da(a.data(), a.size()), db(bt.data(), bt.size()), dc((size_t)m * N);
            check_v(mcv_gemm_create(MCV_F16, m, K, N, da.ptr, db.ptr, 1, dc.ptr, stream, &plan), "hipBLASLt fp16");
            report("hipBLASLt fp16 dense for spmm", [&] { check_v(mcv_run(plan), "mcv_run"); });
            if (func) {
                std::vector<__half> c((size_t)m * N);
                mcb::check_hip(hipMemcpy(c.data(), dc.ptr, c.size() * 2, hipMemcpyDeviceToHost), "copy");
                size_t err = 0;
                for (int i = 0; i < m; ++i)
                    for (int j = 0; j < N; ++j) {
                        double s = 0, mag = 0;
                        for (int t = 0; t < K; ++t) {
                            const double p = (double)__half2float(a[(size_t)i * K + t]) *
                                             __half2float(bt[(size_t)j * K + t]);
                            s += p;
                            mag += std::fabs(p);
                        }
                        err += std::fabs(__half2float(c[(size_t)i * N + j]) - s) > 1e-2 * mag + 1e-3;
                    }
                verdict(err, "hipBLASLt");
            }
        }
    } else if (sparse == 2 || sparse == 3) {
        if (K % V != 0) throw std::runtime_error("blocked baseline needs k to be a multiple of v");
        const bool i8 = sparse == 3;
        const double density = (double)pat.nnz() * V / ((double)m * K);
        std::vector<int> rp, col;
        uniform_blocks(pat.m_vec, K / V, density, gen, rp, col);
        const size_t nnzb = col.size();
        std::printf("A: %d x %d. There are %zu nonzero blocks, Each row has %zu blocks. Block size is %d x %d\n", m,
                    K, nnzb, pat.m_vec ? nnzb / pat.m_vec : 0, V, V);
        std::printf(i8 ? "Blocked int8 SpMM (rocSPARSE BSR)\n" : "Blocked SpMM (rocSPARSE BSR)\n");
        mcb::DeviceBuffer<int> drp(rp.data(), rp.size()), dcol(col.data(), col.size());
        if (i8) {
            const auto vals = int8_values(gen, nnzb * V * V), b = int8_values(gen, (size_t)K * N);
            mcb::DeviceBuffer<int8_t> dv(vals.data(), vals.size()), db(b.data(), b.size());
            mcb::DeviceBuffer<int32_t> dc((size_t)m * N);
            check_v(mcv_bsr_spmm_create(MCV_I8, pat.m_vec, K / V, N, (int64_t)nnzb, V, drp.ptr, dcol.ptr, dv.ptr,
                                        db.ptr, dc.ptr, stream, &plan), "rocSPARSE BSR int8");
            report("rocsparse int8 block bsr", [&] { check_v(mcv_run(plan), "mcv_run"); });
            if (func) {
                std::vector<int32_t> c((size_t)m * N), exp((size_t)m * N, 0);
                mcb::check_hip(hipMemcpy(c.data(), dc.ptr, c.size() * 4, hipMemcpyDeviceToHost), "copy");
                for (int i = 0; i < pat.m_vec; ++i)
                    for (int t = rp[i]; t < rp[i + 1]; ++t)
                        for (int r = 0; r < V; ++r)
                            for (int q = 0; q < V; ++q) {
                                const int av = vals[(size_t)t * V * V + r * V + q];
                                const int8_t *brow = b.data() + (size_t)(col[t] * V + q) * N;
                                int32_t *crow = exp.data() + (size_t)(i * V + r) * N;
                                for (int j = 0; j < N; ++j) crow[j] += av * brow[j];
                            }
                size_t err = 0;
                for (size_t e = 0; e < c.size(); ++e) err += c[e] != exp[e];
                verdict(err, "rocSPARSE BSR");
            }
        } else {
            const auto vals = half_values(gen, nnzb * V * V), b = half_values(gen, (size_t)K * N);
            mcb::DeviceBuffer<__half> dv(vals.data(), vals.size()), db(b.data(), b.size());
            mcb::DeviceBuffer<float> dc((size_t)m * N);
            check_v(mcv_bsr_spmm_create(MCV_F16, pat.m_vec, K / V, N, (int64_t)nnzb, V, drp.ptr, dcol.ptr, dv.ptr,
                                        db.ptr, dc.ptr, stream, &plan), "rocSPARSE BSR fp16");
            report("rocsparse fp16 block bsr", [&] { check_v(mcv_run(plan), "mcv_run"); });
            if (func) {
                std::vector<float> c((size_t)m * N);
                std::vector<double> exp((size_t)m * N, 0.0), mag((size_t)m * N, 0.0);
                mcb::check_hip(hipMemcpy(c.data(), dc.ptr, c.size() * 4, hipMemcpyDeviceToHost), "copy");
                for (int i = 0; i < pat.m_vec; ++i)
                    for (int t = rp[i]; t < rp[i + 1]; ++t)
                        for (int r = 0; r < V; ++r)
                            for (int q = 0; q < V; ++q) {
                                const double av = __half2float(vals[(size_t)t * V * V + r * V + q]);
                                for (int j = 0; j < N; ++j) {
                                    const double p = av * __half2float(b[(size_t)(col[t] * V + q) * N + j]);
                                    exp[(size_t)(i * V + r) * N + j] += p;
                                    mag[(size_t)(i * V + r) * N + j] += std::fabs(p);
                                }
                            }
                size_t err = 0;
                for (size_t e = 0; e < c.size(); ++e) err += std::fabs(c[e] - exp[e]) > 1e-2 * mag[e] + 1e-3;
                verdict(err, "rocSPARSE BSR");
            }
        }
    } else if (sparse == 1 && kernel == 3) {
        if (!mixed) throw std::runtime_error("the CSR baseline is built for half precision (mixed 1)");
        // the vector matrix as scalar CSR: row i*V+v holds vector row i's columns
        std::vector<int> rp((size_t)m + 1), col;
        col.reserve((size_t)pat.nnz() * V);
        for (int i = 0; i < pat.m_vec; ++i)
            for (int v = 0; v < V; ++v) {
                rp[(size_t)i * V + v] = (int)col.size();
                col.insert(col.end(), pat.col_indices.begin() + pat.row_offsets[i],
                           pat.col_indices.begin() + pat.row_offsets[i + 1]);
            }
        rp[m] = (int)col.size();
        std::printf("Using rocSPARSE CSR\n");
        const auto vals = half_values(gen, col.size()), b = half_values(gen, (size_t)K * N);
        mcb::DeviceBuffer<int> drp(rp.data(), rp.size()), dcol(col.data(), col.size());
        mcb::DeviceBuffer<__half> dv(vals.data(), vals.size()), db(b.data(), b.size());
        mcb::DeviceBuffer<float> dc((size_t)m * N);
        check_v(mcv_csr_spmm_create(MCV_F16, MCV_CSR_NNZ_SPLIT, m, K, N, (int64_t)col.size(), drp.ptr, dcol.ptr, dv.ptr,
                                    db.ptr, dc.ptr, stream, &plan), "rocSPARSE CSR fp16");
        report("rocsparse fp16 csr SpMM", [&] { check_v(mcv_run(plan), "mcv_run"); });
        if (func) {
            std::vector<float> c((size_t)m * N);
            mcb::check_hip(hipMemcpy(c.data(), dc.ptr, c.size() * 4, hipMemcpyDeviceToHost), "copy");
            size_t err = 0;
            for (int r = 0; r < m; ++r)
                for (int j = 0; j < N; ++j) {
                    double s = 0, mag = 0;
                    for (int t = rp[r]; t < rp[r + 1]; ++t) {
                        const double p = (double)__half2float(vals[t]) * __half2float(b[(size_t)col[t] * N + j]);
                        s += p;
                        mag += std::fabs(p);
                    }
                    err += std::fabs(c[(size_t)r * N + j] - s) > 1e-2 * mag + 1e-3;
                }
            verdict(err, "rocSPARSE CSR");
        }
    } else {
        std::fprintf(stderr, "sparse %d kernel %d: the reference's own vectorSparse kernels are not carried; "
                             "supported: sparse -1, -2, 2, 3, and 1 with kernel 3\n", sparse, kernel);
        return 2;
    }
    if (plan) check_v(mcv_destroy(plan), "mcv_destroy");
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc != 9) {
        std::printf("usage: ./baseline_spmm_benchmark [bm] [k] [v] [kernel] [sort] [function] [sparse] [mixed]\n");
        std::printf("bm: .smtx path or synthetic:M,K,sparsity[,seed]; k: dense dimension N; sparse: -1 int8 dense, "
                    "-2 fp16 dense, 3 int8 blocked, 2 fp16 blocked, 1 (kernel 3) fp16 CSR\n");
        return argc == 2 ? 0 : 2;
    }
    try {
        std::cout << "Sparse matrix: " << argv[1] << "\n";
        return run(argv[1], std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[6]),
                   std::atoi(argv[7]), std::atoi(argv[8]));
    } catch (const std::exception &e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 2;
    }
}
