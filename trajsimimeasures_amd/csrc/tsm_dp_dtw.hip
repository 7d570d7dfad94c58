// tsm_dp_dtw.hip -- DTW instances of the DP kernel and its launcher.
#include "tsm_launch.hpp"

namespace tsm {
template int run_dp<kDTW, false>(int, const tsm_trajset &, const PairSpace &, int64_t, double, double *, int32_t *, hipStream_t);
template int run_dp<kDTW, true>(int, const tsm_trajset &, const PairSpace &, int64_t, double, double *, int32_t *, hipStream_t);
}  // namespace tsm
