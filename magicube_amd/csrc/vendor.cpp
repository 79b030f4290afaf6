// Vendor-library comparators (libmagicube_vendor.so): rocSPARSE generic SpMM
// (CSR, BSR) and SDDMM, hipBLASLt dense GEMM, behind the plan interface of
// include/magicube_vendor.h.  Host code only; every launch is the library's.
//
// Layout bridge for hipBLASLt (column-major): the row-major product
// C[m x n] = A[m x k] * B[k x n] is computed as the column-major product
// C^T[n x m] = B^T[n x k] * A^T[k x m], in which a row-major matrix is its own
// transpose read column-major.  So the library's first operand is B (op N on
// an n x k view, or op T on a k x n view when B is stored transposed), its
// second is A (op N on a k x m view), and the result is C read as n x m.
#include <hip/hip_runtime_api.h>
#include <hipblaslt/hipblaslt.h>
#include <rocsparse/rocsparse.h>

#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/magicube_vendor.h"

namespace {

thread_local std::string g_err;

mc_status record(mc_status s, const char *what, int code) {
    char buf[320];
    std::snprintf(buf, sizeof(buf), "%s failed with status %d", what, code);
    g_err = buf;
    return s;
}

mc_status from_sparse(rocsparse_status s) {
    switch (s) {
    case rocsparse_status_success: return MC_OK;
    case rocsparse_status_not_implemented: return MC_EUNSUPPORTED;
    case rocsparse_status_invalid_size:
    case rocsparse_status_invalid_pointer:
    case rocsparse_status_invalid_value: return MC_EINVAL_ARG;
    default: return MC_EHIP;
    }
}

mc_status from_lt(hipblasStatus_t s) {
    switch (s) {
    case HIPBLAS_STATUS_SUCCESS: return MC_OK;
    case HIPBLAS_STATUS_NOT_SUPPORTED: return MC_EUNSUPPORTED;
    case HIPBLAS_STATUS_INVALID_VALUE: return MC_EINVAL_ARG;
    default: return MC_EHIP;
    }
}

enum Kind { kSpmm, kSddmm, kGemm };

}  // namespace

struct mcv_plan {
    Kind kind = kSpmm;
    hipStream_t stream = nullptr;
    void *ws = nullptr;
    size_t ws_bytes = 0;

    // rocSPARSE
    rocsparse_handle sp = nullptr;
    rocsparse_spmat_descr mat = nullptr;
    rocsparse_dnmat_descr d0 = nullptr, d1 = nullptr;
    rocsparse_operation op1 = rocsparse_operation_none;
    rocsparse_spmm_alg alg = rocsparse_spmm_alg_default;
    rocsparse_datatype compute = rocsparse_datatype_f32_r;
    int32_t ialpha = 1, ibeta = 0;
    float falpha = 1.0f, fbeta = 0.0f;

    // hipBLASLt
    hipblasLtHandle_t lt = nullptr;
    hipblasLtMatmulDesc_t desc = nullptr;
    hipblasLtMatrixLayout_t la = nullptr, lb = nullptr, lc = nullptr;
    hipblasLtMatmulAlgo_t algo{};
    const void *ga = nullptr, *gb = nullptr;
    void *gc = nullptr;
    bool int_scale = true;

    const void *alpha() const { return compute == rocsparse_datatype_i32_r ? (const void *)&ialpha : (const void *)&falpha; }
    const void *beta() const { return compute == rocsparse_datatype_i32_r ? (const void *)&ibeta : (const void *)&fbeta; }
    const void *lt_alpha() const { return int_scale ? (const void *)&ialpha : (const void *)&falpha; }
    const void *lt_beta() const { return int_scale ? (const void *)&ibeta : (const void *)&fbeta; }
};

namespace {

void release(mcv_plan *p) {
    if (!p) return;
    if (p->stream || p->sp || p->lt) (void)hipStreamSynchronize(p->stream);
    if (p->mat) rocsparse_destroy_spmat_descr(p->mat);
    if (p->d0) rocsparse_destroy_dnmat_descr(p->d0);
    if (p->d1) rocsparse_destroy_dnmat_descr(p->d1);
    if (p->sp) rocsparse_destroy_handle(p->sp);
    if (p->la) hipblasLtMatrixLayoutDestroy(p->la);
    if (p->lb) hipblasLtMatrixLayoutDestroy(p->lb);
    if (p->lc) hipblasLtMatrixLayoutDestroy(p->lc);
    if (p->desc) hipblasLtMatmulDescDestroy(p->desc);
    if (p->lt) hipblasLtDestroy(p->lt);
    if (p->ws) (void)hipFree(p->ws);
    delete p;
}

mc_status alloc_ws(mcv_plan *p, size_t bytes) {
    // rocSPARSE expects a valid pointer even when it asks for no workspace
    p->ws_bytes = bytes;
    const hipError_t e = hipMalloc(&p->ws, bytes ? bytes : 256);
    if (e != hipSuccess) return record(MC_EHIP, "hipMalloc(workspace)", (int)e);
    return MC_OK;
}

#define SP_CHECK(call)                                                               \
    do {                                                                             \
        const rocsparse_status s_ = (call);                                          \
        if (s_ != rocsparse_status_success) {                                        \
            release(p);                                                              \
            return record(from_sparse(s_), #call, (int)s_);                          \
        }                                                                            \
    } while (0)

#define MC_CHECK(call)                                                               \
    do {                                                                             \
        const mc_status s_ = (call);                                                 \
        if (s_ != MC_OK) {                                                           \
            release(p);                                                              \
            return s_;                                                               \
        }                                                                            \
    } while (0)

rocsparse_datatype value_type(mcv_dtype t) {
    return t == MCV_I8 ? rocsparse_datatype_i8_r : rocsparse_datatype_f16_r;
}
rocsparse_datatype acc_type(mcv_dtype t) {
    return t == MCV_I8 ? rocsparse_datatype_i32_r : rocsparse_datatype_f32_r;
}

// Shared tail of the CSR and BSR SpMM plans: dense descriptors, workspace
// query, allocation and the (blocking) preprocess stage.
mc_status finish_spmm(mcv_plan *p, mcv_dtype dtype, int64_t rows, int64_t k, int64_t n, const void *b,
                      void *c) {
    SP_CHECK(rocsparse_create_dnmat_descr(&p->d0, k, n, n, const_cast<void *>(b), value_type(dtype),
                                          rocsparse_order_row));
    SP_CHECK(rocsparse_create_dnmat_descr(&p->d1, rows, n, n, c, acc_type(dtype), rocsparse_order_row));
    size_t bytes = 0;
    SP_CHECK(rocsparse_spmm(p->sp, rocsparse_operation_none, rocsparse_operation_none, p->alpha(), p->mat, p->d0,
                            p->beta(), p->d1, p->compute, p->alg, rocsparse_spmm_stage_buffer_size, &bytes,
                            nullptr));
    MC_CHECK(alloc_ws(p, bytes));
    SP_CHECK(rocsparse_spmm(p->sp, rocsparse_operation_none, rocsparse_operation_none, p->alpha(), p->mat, p->d0,
                            p->beta(), p->d1, p->compute, p->alg, rocsparse_spmm_stage_preprocess, &bytes, p->ws));
    return MC_OK;
}

mc_status new_sparse_plan(mcv_plan **out, Kind kind, rocsparse_datatype compute, hipStream_t stream) {
    mcv_plan *p = new mcv_plan();
    p->kind = kind;
    p->stream = stream;
    p->compute = compute;
    SP_CHECK(rocsparse_create_handle(&p->sp));
    SP_CHECK(rocsparse_set_stream(p->sp, stream));
    *out = p;
    return MC_OK;
}

}  // namespace

extern "C" {

mc_status mcv_csr_spmm_create(mcv_dtype dtype, mcv_csr_alg alg, int64_t m, int64_t k, int64_t n, int64_t nnz,
                              const int *row_ptr, const int *col_ind, const void *values, const void *b, void *c,
                              hipStream_t stream, mcv_plan **plan) {
    if (!plan) return record(MC_EINVAL_ARG, "mcv_csr_spmm_create(plan)", 0);
    *plan = nullptr;
    if (dtype != MCV_I8 && dtype != MCV_F16) return record(MC_EINVAL_ARG, "mcv_csr_spmm_create(dtype)", dtype);
    if (m <= 0 || k <= 0 || n <= 0 || nnz < 0) return record(MC_EINVAL_SHAPE, "mcv_csr_spmm_create(shape)", 0);
    if (!row_ptr || !b || !c || (nnz > 0 && (!col_ind || !values)))
        return record(MC_EINVAL_ARG, "mcv_csr_spmm_create(pointer)", 0);
    rocsparse_spmm_alg a;
    switch (alg) {
    case MCV_CSR_DEFAULT: a = rocsparse_spmm_alg_csr; break;
    case MCV_CSR_ROW_SPLIT: a = rocsparse_spmm_alg_csr_row_split; break;
    case MCV_CSR_NNZ_SPLIT: a = rocsparse_spmm_alg_csr_nnz_split; break;
    case MCV_CSR_MERGE_PATH: a = rocsparse_spmm_alg_csr_merge_path; break;
    default: return record(MC_EINVAL_ARG, "mcv_csr_spmm_create(alg)", alg);
    }
    mcv_plan *p = nullptr;
    const mc_status s = new_sparse_plan(&p, kSpmm, acc_type(dtype), stream);
    if (s != MC_OK) return s;
    p->alg = a;
    SP_CHECK(rocsparse_create_csr_descr(&p->mat, m, k, nnz, const_cast<int *>(row_ptr), const_cast<int *>(col_ind),
                                        const_cast<void *>(values), rocsparse_indextype_i32, rocsparse_indextype_i32,
                                        rocsparse_index_base_zero, value_type(dtype)));
    const mc_status f = finish_spmm(p, dtype, m, k, n, b, c);
    if (f != MC_OK) return f;   // finish_spmm released the plan
    *plan = p;
    return MC_OK;
}

mc_status mcv_bsr_spmm_create(mcv_dtype dtype, int64_t mb, int64_t kb, int64_t n, int64_t nnzb, int block_dim,
                              const int *row_ptr, const int *col_ind, const void *values, const void *b, void *c,
                              hipStream_t stream, mcv_plan **plan) {
    if (!plan) return record(MC_EINVAL_ARG, "mcv_bsr_spmm_create(plan)", 0);
    *plan = nullptr;
    if (dtype != MCV_I8 && dtype != MCV_F16) return record(MC_EINVAL_ARG, "mcv_bsr_spmm_create(dtype)", dtype);
    if (mb <= 0 || kb <= 0 || n <= 0 || nnzb < 0 || block_dim <= 0)
        return record(MC_EINVAL_SHAPE, "mcv_bsr_spmm_create(shape)", 0);
    if (!row_ptr || !b || !c || (nnzb > 0 && (!col_ind || !values)))
        return record(MC_EINVAL_ARG, "mcv_bsr_spmm_create(pointer)", 0);
    mcv_plan *p = nullptr;
    const mc_status s = new_sparse_plan(&p, kSpmm, acc_type(dtype), stream);
    if (s != MC_OK) return s;
    p->alg = rocsparse_spmm_alg_bsr;
    SP_CHECK(rocsparse_create_bsr_descr(&p->mat, mb, kb, nnzb, rocsparse_direction_row, block_dim,
                                        const_cast<int *>(row_ptr), const_cast<int *>(col_ind),
                                        const_cast<void *>(values), rocsparse_indextype_i32, rocsparse_indextype_i32,
                                        rocsparse_index_base_zero, value_type(dtype)));
    const mc_status f = finish_spmm(p, dtype, mb * block_dim, kb * block_dim, n, b, c);
    if (f != MC_OK) return f;
    *plan = p;
    return MC_OK;
}

mc_status mcv_csr_sddmm_create(int64_t m, int64_t n, int64_t k, int64_t nnz, const int *row_ptr, const int *col_ind,
                               const void *lhs, const void *rhs, void *out_values, hipStream_t stream,
                               mcv_plan **plan) {
    if (!plan) return record(MC_EINVAL_ARG, "mcv_csr_sddmm_create(plan)", 0);
    *plan = nullptr;
    if (m <= 0 || n <= 0 || k <= 0 || nnz < 0) return record(MC_EINVAL_SHAPE, "mcv_csr_sddmm_create(shape)", 0);
    if (!row_ptr || !lhs || !rhs || (nnz > 0 && (!col_ind || !out_values)))
        return record(MC_EINVAL_ARG, "mcv_csr_sddmm_create(pointer)", 0);
    mcv_plan *p = nullptr;
    const mc_status s = new_sparse_plan(&p, kSddmm, rocsparse_datatype_f32_r, stream);
    if (s != MC_OK) return s;
    SP_CHECK(rocsparse_create_csr_descr(&p->mat, m, n, nnz, const_cast<int *>(row_ptr), const_cast<int *>(col_ind),
                                        out_values, rocsparse_indextype_i32, rocsparse_indextype_i32,
                                        rocsparse_index_base_zero, rocsparse_datatype_f32_r));
    SP_CHECK(rocsparse_create_dnmat_descr(&p->d0, m, k, k, const_cast<void *>(lhs), rocsparse_datatype_f16_r,
                                          rocsparse_order_row));
    // rhs (n x k row-major) read column-major is the k x n right operand
    // itself; if that order is not implemented, fall back to op T on the
    // row-major n x k view.
    SP_CHECK(rocsparse_create_dnmat_descr(&p->d1, k, n, k, const_cast<void *>(rhs), rocsparse_datatype_f16_r,
                                          rocsparse_order_column));
    size_t bytes = 0;
    rocsparse_status q = rocsparse_sddmm_buffer_size(p->sp, rocsparse_operation_none, rocsparse_operation_none,
                                                     p->alpha(), p->d0, p->d1, p->beta(), p->mat, p->compute,
                                                     rocsparse_sddmm_alg_default, &bytes);
    if (q == rocsparse_status_not_implemented) {
        rocsparse_destroy_dnmat_descr(p->d1);
        p->d1 = nullptr;
        SP_CHECK(rocsparse_create_dnmat_descr(&p->d1, n, k, k, const_cast<void *>(rhs), rocsparse_datatype_f16_r,
                                              rocsparse_order_row));
        p->op1 = rocsparse_operation_transpose;
        q = rocsparse_sddmm_buffer_size(p->sp, rocsparse_operation_none, p->op1, p->alpha(), p->d0, p->d1,
                                        p->beta(), p->mat, p->compute, rocsparse_sddmm_alg_default, &bytes);
    }
    SP_CHECK(q);
    MC_CHECK(alloc_ws(p, bytes));
    SP_CHECK(rocsparse_sddmm_preprocess(p->sp, rocsparse_operation_none, p->op1, p->alpha(), p->d0, p->d1, p->beta(),
                                        p->mat, p->compute, rocsparse_sddmm_alg_default, p->ws));
    *plan = p;
    return MC_OK;
}

mc_status mcv_gemm_create(mcv_dtype dtype, int64_t m, int64_t k, int64_t n, const void *a, const void *b,
                          int b_transposed, void *c, hipStream_t stream, mcv_plan **plan) {
    if (!plan) return record(MC_EINVAL_ARG, "mcv_gemm_create(plan)", 0);
    *plan = nullptr;
    if (dtype != MCV_I8 && dtype != MCV_F16) return record(MC_EINVAL_ARG, "mcv_gemm_create(dtype)", dtype);
    if (m <= 0 || k <= 0 || n <= 0) return record(MC_EINVAL_SHAPE, "mcv_gemm_create(shape)", 0);
    if (!a || !b || !c) return record(MC_EINVAL_ARG, "mcv_gemm_create(pointer)", 0);
    const bool i8 = dtype == MCV_I8;
    const hipDataType ab = i8 ? HIP_R_8I : HIP_R_16F;
    const hipDataType cd = i8 ? HIP_R_32I : HIP_R_16F;
    const hipblasComputeType_t ct = i8 ? HIPBLAS_COMPUTE_32I : HIPBLAS_COMPUTE_32F;
    const int32_t opa = b_transposed ? HIPBLAS_OP_T : HIPBLAS_OP_N, opb = HIPBLAS_OP_N;

    mcv_plan *p = new mcv_plan();
    p->kind = kGemm;
    p->stream = stream;
    p->ga = b;
    p->gb = a;
    p->gc = c;
#define LT_CHECK(call)                                                               \
    do {                                                                             \
        const hipblasStatus_t s_ = (call);                                           \
        if (s_ != HIPBLAS_STATUS_SUCCESS) {                                          \
            release(p);                                                              \
            return record(from_lt(s_), #call, (int)s_);                              \
        }                                                                            \
    } while (0)
    LT_CHECK(hipblasLtCreate(&p->lt));
    if (b_transposed) LT_CHECK(hipblasLtMatrixLayoutCreate(&p->la, ab, k, n, k));
    else LT_CHECK(hipblasLtMatrixLayoutCreate(&p->la, ab, n, k, n));
    LT_CHECK(hipblasLtMatrixLayoutCreate(&p->lb, ab, k, m, k));
    LT_CHECK(hipblasLtMatrixLayoutCreate(&p->lc, cd, n, m, n));

    hipblasLtMatmulPreference_t pref = nullptr;
    LT_CHECK(hipblasLtMatmulPreferenceCreate(&pref));
    const uint64_t max_ws = 64ull << 20;
    hipblasLtMatmulPreferenceSetAttribute(pref, HIPBLASLT_MATMUL_PREF_MAX_WORKSPACE_BYTES, &max_ws, sizeof(max_ws));

    // Integer GEMMs scale in int32 where the library offers it, else fp32.
    hipblasLtMatmulHeuristicResult_t res[8];
    int got = 0;
    const int scale_tries = i8 ? 2 : 1;
    for (int t = 0; t < scale_tries && got == 0; ++t) {
        p->int_scale = i8 && t == 0;
        if (p->desc) {
            hipblasLtMatmulDescDestroy(p->desc);
            p->desc = nullptr;
        }
        if (hipblasLtMatmulDescCreate(&p->desc, ct, p->int_scale ? HIP_R_32I : HIP_R_32F) != HIPBLAS_STATUS_SUCCESS)
            continue;
        hipblasLtMatmulDescSetAttribute(p->desc, HIPBLASLT_MATMUL_DESC_TRANSA, &opa, sizeof(opa));
        hipblasLtMatmulDescSetAttribute(p->desc, HIPBLASLT_MATMUL_DESC_TRANSB, &opb, sizeof(opb));
        if (hipblasLtMatmulAlgoGetHeuristic(p->lt, p->desc, p->la, p->lb, p->lc, p->lc, pref, 8, res, &got) !=
            HIPBLAS_STATUS_SUCCESS)
            got = 0;
    }
    hipblasLtMatmulPreferenceDestroy(pref);
    if (got == 0) {
        release(p);
        return record(MC_EUNSUPPORTED, "hipblasLtMatmulAlgoGetHeuristic (no algorithm)", 0);
    }
    size_t need = 0;
    for (int i = 0; i < got; ++i)
        if (res[i].state == HIPBLAS_STATUS_SUCCESS && res[i].workspaceSize > need) need = res[i].workspaceSize;
    MC_CHECK(alloc_ws(p, need));

    // Keep the fastest candidate: 2 untimed and 5 timed launches each.
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        if (e0) (void)hipEventDestroy(e0);
        release(p);
        return record(MC_EHIP, "hipEventCreate", 0);
    }
    float best = -1.0f;
    for (int i = 0; i < got; ++i) {
        if (res[i].state != HIPBLAS_STATUS_SUCCESS) continue;
        bool ok = true;
        for (int r = 0; r < 2 && ok; ++r)
            ok = hipblasLtMatmul(p->lt, p->desc, p->lt_alpha(), p->ga, p->la, p->gb, p->lb, p->lt_beta(), p->gc,
                                 p->lc, p->gc, p->lc, &res[i].algo, p->ws, p->ws_bytes, stream) ==
                 HIPBLAS_STATUS_SUCCESS;
        if (!ok) continue;
        (void)hipEventRecord(e0, stream);
        for (int r = 0; r < 5 && ok; ++r)
            ok = hipblasLtMatmul(p->lt, p->desc, p->lt_alpha(), p->ga, p->la, p->gb, p->lb, p->lt_beta(), p->gc,
                                 p->lc, p->gc, p->lc, &res[i].algo, p->ws, p->ws_bytes, stream) ==
                 HIPBLAS_STATUS_SUCCESS;
        (void)hipEventRecord(e1, stream);
        (void)hipEventSynchronize(e1);
        float ms = 0.0f;
        if (!ok || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) continue;
        if (best < 0.0f || ms < best) {
            best = ms;
            p->algo = res[i].algo;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (best < 0.0f) {
        release(p);
        return record(MC_EUNSUPPORTED, "hipblasLtMatmul (no candidate ran)", 0);
    }
#undef LT_CHECK
    *plan = p;
    return MC_OK;
}

mc_status mcv_run(mcv_plan *p) {
    if (!p) return record(MC_EINVAL_ARG, "mcv_run(plan)", 0);
    size_t bytes = p->ws_bytes;
    switch (p->kind) {
    case kSpmm: {
        const rocsparse_status s =
            rocsparse_spmm(p->sp, rocsparse_operation_none, rocsparse_operation_none, p->alpha(), p->mat, p->d0,
                           p->beta(), p->d1, p->compute, p->alg, rocsparse_spmm_stage_compute, &bytes, p->ws);
        return s == rocsparse_status_success ? MC_OK : record(from_sparse(s), "rocsparse_spmm", (int)s);
    }
    case kSddmm: {
        const rocsparse_status s = rocsparse_sddmm(p->sp, rocsparse_operation_none, p->op1, p->alpha(), p->d0, p->d1,
                                                   p->beta(), p->mat, p->compute, rocsparse_sddmm_alg_default, p->ws);
        return s == rocsparse_status_success ? MC_OK : record(from_sparse(s), "rocsparse_sddmm", (int)s);
    }
    case kGemm: {
        const hipblasStatus_t s =
            hipblasLtMatmul(p->lt, p->desc, p->lt_alpha(), p->ga, p->la, p->gb, p->lb, p->lt_beta(), p->gc, p->lc,
                            p->gc, p->lc, &p->algo, p->ws, p->ws_bytes, p->stream);
        return s == HIPBLAS_STATUS_SUCCESS ? MC_OK : record(from_lt(s), "hipblasLtMatmul", (int)s);
    }
    }
    return record(MC_EINVAL_ARG, "mcv_run(kind)", (int)p->kind);
}

size_t mcv_workspace_bytes(const mcv_plan *p) { return p ? p->ws_bytes : 0; }

mc_status mcv_destroy(mcv_plan *p) {
    release(p);
    return MC_OK;
}

const char *mcv_last_error(void) { return g_err.c_str(); }

}  // extern "C"
