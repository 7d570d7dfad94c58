/*
 * tsm_cluster.h -- C ABI of the clustering step over the all-pairs matrix
 * (libtsm.so).
 *
 * The reference's clustering benchmark builds the symmetric all-pairs
 * matrix from the upper triangle (test_clustering_time.py:219,256-269) and
 * clusters it with kmedoids.fasterpam(dists, Config.clustering_num_centroids)
 * (test_clustering_time.py:353-367, config.py:72; the kmedoids package,
 * kmedoids==0.4.3, requirements377.txt:6).  These entry points keep the
 * matrix on the device: tsm_square_from_triu expands a tsm_pairs_triu result
 * in place of the host symmetrisation, and tsm_kmedoids_fasterpam runs
 * FasterPAM (Schubert & Rousseeuw, Information Systems 101, 2021) on it.
 * Conventions as in tsm.h: TSM_OK or a negative code, tsm_last_error(),
 * device pointers on the current device, work ordered on `stream`.
 */
#ifndef TSM_CLUSTER_H
#define TSM_CLUSTER_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/*
 * Symmetric [n][n] matrix (row-major, zero diagonal) from the row-major
 * strict upper triangle tri [n*(n-1)/2] (itertools.combinations order):
 * replaces test_clustering_time.py:256-269.  diss: device float64 [n*n].
 */
int tsm_square_from_triu(const double *tri, int64_t n, double *diss, void *stream);

/*
 * FasterPAM k-medoids on the device dissimilarity matrix diss [n][n]
 * (float64, row-major; the reference passes the symmetric all-pairs
 * matrix; no NaN).  Replaces kmedoids.fasterpam(diss, medoids, max_iter)
 * (test_clustering_time.py:358).
 *   medoids  host int64 [k]: the initial medoids (distinct indices in
 *            [0, n)) on entry, the final ones on return.  1 <= k <= 64.
 *   labels   device int32 [n] or NULL: each point's medoid, as a position in
 *            medoids (kmedoids' labels).
 *   loss, n_iter, n_swap  host outputs or NULL: the sum of distances to the
 *            nearest medoid, the passes run and the swaps made.
 * Candidates are scanned in index order and a swap is made as soon as one
 * lowers the loss; a pass ends at the candidate of the last swap; passes
 * repeat until one makes no swap or max_iter passes ran.  The function
 * synchronises `stream` (it reads each batch's outcome on the host).
 */
int tsm_kmedoids_fasterpam(const double *diss, int64_t n, int32_t k, int64_t *medoids, int32_t max_iter,
                           int32_t *labels, double *loss, int32_t *n_iter, int32_t *n_swap, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* TSM_CLUSTER_H */
