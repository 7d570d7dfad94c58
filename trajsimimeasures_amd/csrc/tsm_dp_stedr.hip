// tsm_dp_stedr.hip -- STEDR instances of the DP kernel (FP64 coordinates and
// timestamps) and its launcher.
#include "tsm_launch.hpp"

namespace tsm {
template int run_dp<kSTEDR, true>(Workspace &, int, const tsm_trajset &, const PairSpace &, int64_t, double, double *, int32_t *, hipStream_t, double, double, double);
}  // namespace tsm
