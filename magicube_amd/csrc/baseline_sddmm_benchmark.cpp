// baseline_sddmm_benchmark: the command line and report lines of the
// reference's vendor-library SDDMM baselines (baselines/sddmm_benchmark.cpp:
// 480-540 for argv, :33-88 cuSPARSE SDDMM, :340-430 dense GEMM), driving
// rocSPARSE and hipBLASLt through libmagicube_vendor.so.
//
//   baseline_sddmm_benchmark bm k v kernel alg sort function sparse mixed
//
// bm is a DLMC .smtx file or "synthetic:M,N,sparsity[,seed]" (the mask, M
// rows of V-row blocks); k is the inner dimension.  Modes (the reference's
// launch_sddmm_*.py settings):
//   sparse  0          fp16 dense product lhs (M x k) * rhs^T (N x k) (hipBLASLt)
//   sparse -1          the same in int8 (int32 output)
//   sparse  1 kernel 3 fp16 SDDMM over the mask as scalar CSR (rocSPARSE;
//                      fp32 output values)
// The reference's own vectorSparse SDDMM kernels are not carried.  Timing,
// values and verification as in baseline_spmm_benchmark.
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

template <typename F> void report(const char *label, F launch) {
    const float ms = mcb::time_launches(launch, 16, 512);
    std::cout << label << " runtime " << ms << " ms\n";
}

void verdict(size_t errors, const char *who) {
    if (errors) std::printf("%s does not agree with SEQUENTIAL! %zu errors!\n", who, errors);
    else std::printf("Results verified: they agree.\n");
}

int run(const std::string &bm, int K, int V, int kernel, int func, int sparse, int mixed) {
    const Pattern pat = mcb::load_pattern(bm, V);
    const int M = pat.m_vec * V, N = pat.k;
    std::printf("Problem size: M: %d, N: %d, nnz: %d, K: %d\n", M, N, pat.nnz() * V, K);
    std::default_random_engine gen;
    hipStream_t stream = nullptr;
    mcv_plan *plan = nullptr;

    if (sparse == 0 || sparse == -1) {
        const bool i8 = sparse == -1;
        std::printf(i8 ? "int8 Dense Baseline\n" : "f16 Dense Baseline\n");
        if (i8) {
            const auto a = int8_values(gen, (size_t)M * K), b = int8_values(gen, (size_t)N * K);
            mcb::DeviceBuffer<int8_t> da(a.data(), a.size()), db(b.data(), b.size());
            mcb::DeviceBuffer<int32_t> dc((size_t)M * N);
            check_v(mcv_gemm_create(MCV_I8, M, K, N, da.ptr, db.ptr, 1, dc.ptr, stream, &plan), "hipBLASLt int8");
            report("Int8 sddmm dense (hipBLASLt)", [&] { check_v(mcv_run(plan), "mcv_run"); });
            if (func) {
                std::vector<int32_t> c((size_t)M * N);
                mcb::check_hip(hipMemcpy(c.data(), dc.ptr, c.size() * 4, hipMemcpyDeviceToHost), "copy");
                size_t err = 0;
                for (int i = 0; i < M; ++i)
                    for (int j = 0; j < N; ++j) {
                        long long s = 0;
                        for (int t = 0; t < K; ++t) s += (long long)a[(size_t)i * K + t] * b[(size_t)j * K + t];
                        err += (int32_t)s != c[(size_t)i * N + j];
                    }
                verdict(err, "hipBLASLt");
            }
        } else {
            const auto a = half_values(gen, (size_t)M * K), b = half_values(gen, (size_t)N * K);
            mcb::DeviceBuffer<__half> da(a.data(), a.size()), db(b.data(), b.size()), dc((size_t)M * N);
            check_v(mcv_gemm_create(MCV_F16, M, K, N, da.ptr, db.ptr, 1, dc.ptr, stream, &plan), "hipBLASLt fp16");
            report("F16 sddmm dense (hipBLASLt)", [&] { check_v(mcv_run(plan), "mcv_run"); });
            if (func) {
                std::vector<__half> c((size_t)M * N);
                mcb::check_hip(hipMemcpy(c.data(), dc.ptr, c.size() * 2, hipMemcpyDeviceToHost), "copy");
                size_t err = 0;
                for (int i = 0; i < M; ++i)
                    for (int j = 0; j < N; ++j) {
                        double s = 0, mag = 0;
                        for (int t = 0; t < K; ++t) {
                            const double p = (double)__half2float(a[(size_t)i * K + t]) * __half2float(b[(size_t)j * K + t]);
                            s += p;
                            mag += std::fabs(p);
                        }
                        err += std::fabs(__half2float(c[(size_t)i * N + j]) - s) > 1e-2 * mag + 1e-3;
                    }
                verdict(err, "hipBLASLt");
            }
        }
    } else if (sparse == 1 && kernel == 3) {
        if (!mixed) throw std::runtime_error("the SDDMM baseline is built for half precision (mixed 1)");
        std::vector<int> rp((size_t)M + 1), col;
        col.reserve((size_t)pat.nnz() * V);
        for (int i = 0; i < pat.m_vec; ++i)
            for (int v = 0; v < V; ++v) {
                rp[(size_t)i * V + v] = (int)col.size();
                col.insert(col.end(), pat.col_indices.begin() + pat.row_offsets[i],
                           pat.col_indices.begin() + pat.row_offsets[i + 1]);
            }
        rp[M] = (int)col.size();
        std::printf("Using rocSPARSE\n");
        const auto a = half_values(gen, (size_t)M * K), b = half_values(gen, (size_t)N * K);
        mcb::DeviceBuffer<int> drp(rp.data(), rp.size()), dcol(col.data(), col.size());
        mcb::DeviceBuffer<__half> da(a.data(), a.size()), db(b.data(), b.size());
        std::vector<float> ones(col.size(), 1.0f);
        mcb::DeviceBuffer<float> dout(ones.data(), ones.size());
        check_v(mcv_csr_sddmm_create(M, N, K, (int64_t)col.size(), drp.ptr, dcol.ptr, da.ptr, db.ptr, dout.ptr, stream,
                                     &plan), "rocSPARSE SDDMM");
        report("rocSPARSE SDDMM", [&] { check_v(mcv_run(plan), "mcv_run"); });
        if (func) {
            std::vector<float> c(col.size());
            mcb::check_hip(hipMemcpy(c.data(), dout.ptr, c.size() * 4, hipMemcpyDeviceToHost), "copy");
            size_t err = 0;
            for (int r = 0; r < M; ++r)
                for (int t = rp[r]; t < rp[r + 1]; ++t) {
                    double s = 0, mag = 0;
                    for (int q = 0; q < K; ++q) {
                        const double p = (double)__half2float(a[(size_t)r * K + q]) * __half2float(b[(size_t)col[t] * K + q]);
                        s += p;
                        mag += std::fabs(p);
                    }
                    err += std::fabs(c[t] - s) > 1e-2 * mag + 1e-3;
                }
            verdict(err, "rocSPARSE SDDMM");
        }
    } else {
        std::fprintf(stderr, "sparse %d kernel %d: the reference's own vectorSparse kernels are not carried; "
                             "supported: sparse 0, -1, and 1 with kernel 3\n", sparse, kernel);
        return 2;
    }
    if (plan) check_v(mcv_destroy(plan), "mcv_destroy");
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc != 10) {
        std::printf("usage: ./baseline_sddmm_benchmark [bm] [k] [v] [kernel] [alg] [sort] [function] [sparse] "
                    "[mixed]\n");
        std::printf("bm: .smtx path or synthetic:M,N,sparsity[,seed]; k: inner dimension; sparse: 0 fp16 dense, "
                    "-1 int8 dense, 1 (kernel 3) fp16 rocSPARSE SDDMM\n");
        return argc == 2 ? 0 : 2;
    }
    try {
        std::cout << "Sparse matrix: " << argv[1] << "\n";
        return run(argv[1], std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[7]),
                   std::atoi(argv[8]), std::atoi(argv[9]));
    } catch (const std::exception &e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 2;
    }
}
