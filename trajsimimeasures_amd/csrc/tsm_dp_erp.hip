// tsm_dp_erp.hip -- ERP instances of the DP kernel and its launcher, FP32 DP.
#include "tsm_launch.hpp"

namespace tsm {
template int run_dp<kERP, false>(Workspace &, int, const tsm_trajset &, const PairSpace &, int64_t, double, double *, int32_t *, hipStream_t, double, double, double);
}  // namespace tsm
