// tsm_dp_edr.hip -- EDR instances of the DP kernel (FP32 match test with a
// guard band) and its launcher.
#include "tsm_launch.hpp"

namespace tsm {
template int run_dp<kEDR, false>(Workspace &, int, const tsm_trajset &, const PairSpace &, int64_t, double, double *, int32_t *, hipStream_t, double, double, double);
}  // namespace tsm
