// tsm_dp_dtw64.hip -- DTW instances of the DP kernel and its launcher, FP64 DP (drift-check rerun).
#include "tsm_launch.hpp"

namespace tsm {
template int run_dp<kDTW, true>(Workspace &, int, const tsm_trajset &, const PairSpace &, int64_t, double, double *, int32_t *, hipStream_t, double, double, double);
}  // namespace tsm
