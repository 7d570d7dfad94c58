// tsm_dp_frechet64.hip -- discrete Frechet instances of the DP kernel and its launcher, FP64 DP (drift-check rerun).
#include "tsm_launch.hpp"

namespace tsm {
template int run_dp<kFrechet, true>(Workspace &, int, const tsm_trajset &, const PairSpace &, int64_t, double, double *, int32_t *, hipStream_t, double, double, double);
}  // namespace tsm
