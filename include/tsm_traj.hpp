/*
 * tsm_traj.hpp -- C++ adapter for the reference's C++ kNN indexes: the body
 * of Traj::TrajSimiGPU (query/cpp/spatial/traj.cpp:242-360) as one libtsm.so
 * call through the C ABI (tsm.h) instead of an embedded CPython calling
 * core.*.  DITA and DFT call it in their sampling and verify steps with
 * candidate pairs (query/cpp/DITA.cpp:347,403; query/cpp/DFT.cpp:147,197).
 *
 * TrajPtr is any pointer-like type whose target has NumPoints() and
 * points[j]->x / points[j]->y, as the reference's Traj and Point do
 * (query/cpp/spatial/traj.h, point.h).  Measure names are the reference's:
 * "DTW", "ERP", "Hausdorff", "Frechet" (traj.cpp:257-272), plus "EDR" and
 * "LCSS" with `eps`.  ERP uses the gap point (0, 0), as the reference passes
 * (traj.cpp:314-318).
 *
 * Differences from the reference body: dists receives the value of pair i
 * at position i (the reference appends pair 1's value n times,
 * traj.cpp:336-338), and errors come back as the TSM_E* code with
 * tsm_last_error() instead of a message on stdout.
 */
#ifndef TSM_TRAJ_HPP
#define TSM_TRAJ_HPP

#include <cstdint>
#include <string>
#include <vector>

#include "tsm.h"

template <typename TrajPtr>
inline int tsm_traj_simi(const std::vector<TrajPtr> &trajs1, const std::vector<TrajPtr> &trajs2,
                         const std::string &measure, std::vector<float> &dists, double eps = 0.0) {
    int m;
    if (measure == "DTW") m = TSM_DTW;
    else if (measure == "ERP") m = TSM_ERP;
    else if (measure == "Hausdorff") m = TSM_HAUSDORFF;
    else if (measure == "Frechet") m = TSM_FRECHET;
    else if (measure == "EDR") m = TSM_EDR;
    else if (measure == "LCSS") m = TSM_LCSS;
    else return TSM_E_ARG;
    if (trajs1.size() != trajs2.size()) return TSM_E_ARG;
    const int64_t n = (int64_t)trajs1.size();
    if (n == 0) return TSM_OK;
    // pack trajs1 then trajs2 into one set: offsets [2n + 1], xy [npts][2]
    std::vector<int64_t> offs(2 * n + 1, 0);
    for (int64_t i = 0; i < 2 * n; ++i) {
        const auto &t = i < n ? trajs1[i] : trajs2[i - n];
        offs[i + 1] = offs[i] + (int64_t)t->NumPoints();
    }
    std::vector<double> xy(2 * offs[2 * n]);
    for (int64_t i = 0; i < 2 * n; ++i) {
        const auto &t = i < n ? trajs1[i] : trajs2[i - n];
        for (int64_t j = 0; j < (int64_t)t->NumPoints(); ++j) {
            xy[2 * (offs[i] + j)] = (double)t->points[j]->x;
            xy[2 * (offs[i] + j) + 1] = (double)t->points[j]->y;
        }
    }
    std::vector<int32_t> pa(n), pb(n);
    for (int64_t i = 0; i < n; ++i) {
        pa[i] = (int32_t)i;
        pb[i] = (int32_t)(n + i);
    }
    tsm_params prm;
    tsm_params_default(&prm);
    prm.eps = eps;
    std::vector<double> out(n);
    const int rc = tsm_pairs_host(m, xy.data(), nullptr, offs.data(), 2 * n, offs[2 * n], pa.data(), pb.data(), n,
                                  &prm, out.data(), nullptr);
    if (rc != TSM_OK) return rc;
    for (int64_t i = 0; i < n; ++i) dists.push_back((float)out[i]);
    return TSM_OK;
}

#endif /* TSM_TRAJ_HPP */
