// SpMM instantiations for the L16-R16 precision pair.
#include "spmm_dispatch.hpp"

namespace mc {
mc_status spmm_launch_l16r16(bool sgn, int V, int cfg, const SpmmArgs &a, hipStream_t s) {
    return spmm_launch_pair<16, 16>(sgn, V, cfg, a, s);
}
}  // namespace mc
