// tsm_dp_stlcss.hip -- STLCSS instances of the DP kernel (FP64 coordinates and
// timestamps) and its launcher.
#include "tsm_launch.hpp"

namespace tsm {
template int run_dp<kSTLCSS, true>(Workspace &, int, const tsm_trajset &, const PairSpace &, int64_t, double, double *, int32_t *, hipStream_t, double, double, double);
}  // namespace tsm
