// Library metadata and error reporting for the C ABI.
#include "mc_internal.hpp"

namespace {
thread_local int g_last_hip_error = 0;
}

namespace mc {
mc_status hip_fail(hipError_t e) {
    g_last_hip_error = (int)e;
    return MC_EHIP;
}
}  // namespace mc

extern "C" {

int mc_version(void) { return 1; }

int mc_last_hip_error(void) { return g_last_hip_error; }

const char *mc_status_string(int s) {
    switch (s) {
    case MC_OK: return "ok";
    case MC_EINVAL_V: return "vec_length must be 2, 4 or 8";
    case MC_EINVAL_SHAPE: return "shape out of range or misaligned";
    case MC_EINVAL_ARG: return "invalid argument";
    case MC_EHIP: return "HIP runtime error";
    case MC_ENOMEM: return "workspace too small";
    case MC_EUNSUPPORTED: return "unsupported combination";
    default: return "unknown status";
    }
}

}  // extern "C"
