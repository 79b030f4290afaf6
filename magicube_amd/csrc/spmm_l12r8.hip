// SpMM instantiations for the L12-R8 precision pair.
#include "spmm_dispatch.hpp"

namespace mc {
mc_status spmm_launch_l12r8(bool sgn, int V, int cfg, const SpmmArgs &a, hipStream_t s) {
    return spmm_launch_pair<12, 8>(sgn, V, cfg, a, s);
}
}  // namespace mc
