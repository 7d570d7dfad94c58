// tsm_dp_frechet.hip -- discrete Frechet instances of the DP kernel and its launcher, FP32 DP.
#include "tsm_launch.hpp"

namespace tsm {
template int run_dp<kFrechet, false>(Workspace &, int, const tsm_trajset &, const PairSpace &, int64_t, double, double *, int32_t *, hipStream_t, double, double, double);
}  // namespace tsm
