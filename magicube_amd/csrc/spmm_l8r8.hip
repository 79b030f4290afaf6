// SpMM instantiations for the L8-R8 precision pair.
#include "spmm_dispatch.hpp"

namespace mc {
mc_status spmm_launch_l8r8(bool sgn, int V, int cfg, const SpmmArgs &a, hipStream_t s) {
    return spmm_launch_pair<8, 8>(sgn, V, cfg, a, s);
}
}  // namespace mc
