// tsm_dp_lcss64.hip -- LCSS instances with the match test on FP64 coordinates
// (sets without a usable coordinate bound, or precision = FP64).
#include "tsm_launch.hpp"

namespace tsm {
template int run_dp<kLCSS, true>(Workspace &, int, const tsm_trajset &, const PairSpace &, int64_t, double, double *, int32_t *, hipStream_t, double, double, double);
}  // namespace tsm
