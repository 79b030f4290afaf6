/*
 * magicube_compat.hpp — C++ shim with the reference's entry-point names.
 *
 * Callers of Magicube's
 *   spmm::wmmaSpmm_{4b,8b4b,12b4b,16b4b,8b,12b8b,16b8b,16b}
 *                                       (SpMM/SpMM/include/wmma_spmm.cuh:15-78)
 *   sddmm::wmmaSddmm_{4b,8b,16b}        (SDDMM/SDDMM/include/wmma_sddmm.cuh:7-32)
 * switch by including this header instead of the reference's and linking
 * libmagicube_hip.so.  Arguments and their order are unchanged.  The trailing
 * stream (default: null stream) and lane interpretation (default: unsigned,
 * which is what the reference's u8/u4 MMAs compute) are optional.
 *
 * The return value is a hipError_t: hipSuccess, hipErrorInvalidValue for
 * rejected arguments (the reference validates nothing and returns a stale
 * error), hipErrorNotSupported for a combination that is not built, or the
 * HIP error of a failed launch.
 */
#ifndef MAGICUBE_COMPAT_HPP
#define MAGICUBE_COMPAT_HPP

#include <hip/hip_runtime_api.h>

#include "magicube_hip.h"

namespace magicube {

inline hipError_t to_hip(mc_status s) {
    switch (s) {
    case MC_OK: return hipSuccess;
    case MC_EHIP: return static_cast<hipError_t>(mc_last_hip_error());
    case MC_EUNSUPPORTED: return hipErrorNotSupported;
    case MC_ENOMEM: return hipErrorOutOfMemory;
    default: return hipErrorInvalidValue;
    }
}

}  // namespace magicube

namespace spmm {

#define MAGICUBE_SPMM_SHIM(name, fn)                                                                      \
    inline hipError_t name(int m_vec, int vec_length, int n, int k, const int *row_indices,              \
                           const int *row_offsets, const int *column_indices, const int *values,         \
                           const int *rhs_matrix, int *output_matrix, hipStream_t stream = nullptr,      \
                           mc_signedness sign = MC_UNSIGNED) {                                            \
        return magicube::to_hip(fn(m_vec, vec_length, n, k, row_indices, row_offsets, column_indices,    \
                                   values, rhs_matrix, output_matrix, sign, stream));                    \
    }
MAGICUBE_SPMM_SHIM(wmmaSpmm_8b, mc_spmm_l8r8)
MAGICUBE_SPMM_SHIM(wmmaSpmm_16b8b, mc_spmm_l16r8)
MAGICUBE_SPMM_SHIM(wmmaSpmm_8b4b, mc_spmm_l8r4)
MAGICUBE_SPMM_SHIM(wmmaSpmm_4b, mc_spmm_l4r4)
MAGICUBE_SPMM_SHIM(wmmaSpmm_12b8b, mc_spmm_l12r8)
MAGICUBE_SPMM_SHIM(wmmaSpmm_16b, mc_spmm_l16r16)
MAGICUBE_SPMM_SHIM(wmmaSpmm_12b4b, mc_spmm_l12r4)
MAGICUBE_SPMM_SHIM(wmmaSpmm_16b4b, mc_spmm_l16r4)
#undef MAGICUBE_SPMM_SHIM

}  // namespace spmm

namespace sddmm {

#define MAGICUBE_SDDMM_SHIM(name, fn)                                                                     \
    inline hipError_t name(int m_vec, int k, int n, const int *row_indices, const int *row_offsets,      \
                           const int *col_indices, const int *lhs_matrix, const int *rhs_matrix,         \
                           int *output_values, int vec_length, hipStream_t stream = nullptr,             \
                           mc_signedness sign = MC_UNSIGNED) {                                            \
        return magicube::to_hip(fn(m_vec, k, n, row_indices, row_offsets, col_indices, lhs_matrix,       \
                                   rhs_matrix, output_values, vec_length, sign, stream));                \
    }
MAGICUBE_SDDMM_SHIM(wmmaSddmm_8b, mc_sddmm_l8r8)
MAGICUBE_SDDMM_SHIM(wmmaSddmm_4b, mc_sddmm_l4r4)
MAGICUBE_SDDMM_SHIM(wmmaSddmm_16b, mc_sddmm_l16r16)
#undef MAGICUBE_SDDMM_SHIM

}  // namespace sddmm

#endif  // MAGICUBE_COMPAT_HPP
