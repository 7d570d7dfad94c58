// tsm_dp_dtw.hip -- DTW instances of the DP kernel and its launcher, FP32 DP.
#include "tsm_launch.hpp"

namespace tsm {
template int run_dp<kDTW, false>(Workspace &, int, const tsm_trajset &, const PairSpace &, int64_t, double, double *, int32_t *, hipStream_t, double, double, double);
}  // namespace tsm
