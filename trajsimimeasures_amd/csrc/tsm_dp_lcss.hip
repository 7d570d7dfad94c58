// tsm_dp_lcss.hip -- LCSS instances of the DP kernel and its launcher.
#include "tsm_launch.hpp"

namespace tsm {
template int run_dp<kLCSS, false>(Workspace &, int, const tsm_trajset &, const PairSpace &, int64_t, double, double *, int32_t *, hipStream_t, double, double, double);
}  // namespace tsm
