// Host-side helpers shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/magicube_hip.h"

namespace mc {

// Records a HIP failure for mc_last_hip_error() and maps it to MC_EHIP.
mc_status hip_fail(hipError_t e);

inline mc_status check_launch() {
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? MC_OK : hip_fail(e);
}

inline bool valid_v(int v) { return v == 2 || v == 4 || v == 8; }

}  // namespace mc
