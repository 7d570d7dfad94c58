// tsm_dp_erp64.hip -- ERP instances of the DP kernel and its launcher, FP64 DP (drift-check rerun).
#include "tsm_launch.hpp"

namespace tsm {
template int run_dp<kERP, true>(Workspace &, int, const tsm_trajset &, const PairSpace &, int64_t, double, double *, int32_t *, hipStream_t, double, double, double);
}  // namespace tsm
