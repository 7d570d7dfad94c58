/*
 * tsm.h -- C ABI of libtsm.so, the MI355X (gfx950) engine for batched
 * trajectory-similarity dynamic programs: DTW, EDR, LCSS, ERP, discrete
 * Frechet and Hausdorff.
 *
 * The reference exposes one Python function per measure over zero-padded
 * batches (core/dtw.py:56, core/edr.py:55, core/lcss.py:56, core/erp.py:94,
 * core/discret_frechet.py:53, core/hausdorff.py:97), each of which copies
 * the batch to the device, launches one Numba kernel and copies the
 * distances back.  This ABI replaces those bodies: trajectories travel as a
 * packed point buffer plus an offsets array, pairs as index arrays, and one
 * call computes a whole batch on one stream.  Python binds it with ctypes
 * (trajsimimeasures_amd/_lib.py).
 *
 * Conventions
 *   - Every function returns TSM_OK (0) or a negative TSM_E* code; the
 *     message of the last failure on the calling thread is tsm_last_error().
 *   - "device" pointers are HIP device memory on the current device;
 *     "host" pointers are ordinary host memory.
 *   - Device-pointer entry points are stream-ordered on `stream`
 *     (a hipStream_t, NULL = the null stream) and do not synchronise, except
 *     when the library's workspace has to grow (a hipMalloc).  Calls share
 *     one device workspace; a call on a different stream than the previous
 *     call first makes its stream wait for that call's kernels, so calls
 *     from several streams or threads are safe and run one after another.
 *     A captured graph (tsm_graph_capture_pairs) owns its own workspace.
 *   - Results are float64 like the reference's (core/dtw.py:70,76); EDR and
 *     LCSS also report the integer DP count.
 *   - Trajectory lengths must be >= 1.
 *   - A pair whose trajectory index lies outside [0, n) is not computed: its
 *     result is NaN (count -1); no memory outside the set is read.  So is a
 *     pair of the DP measures longer than the set's max_len admits (a
 *     max_len below a trajectory's length): NaN, count -1.
 *   - Coordinates and timestamps must be finite; the host entry points
 *     reject others with TSM_E_ARG.
 *   - SAR uses eps as distance_eps (a sample pair matches when its distance
 *     is <= eps) and time_step / target_length / replicas; timestamps must
 *     not be NaN.
 */
#ifndef TSM_H
#define TSM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TSM_ABI_VERSION 4

/* Measure ids. */
#define TSM_DTW 0       /* core/dtw.py:10-77 */
#define TSM_EDR 1       /* core/edr.py:10-76 */
#define TSM_LCSS 2      /* core/lcss.py:12-77 */
#define TSM_ERP 3       /* core/erp.py:10-120 */
#define TSM_FRECHET 4   /* core/discret_frechet.py:9-74 */
#define TSM_HAUSDORFF 5 /* core/hausdorff.py:12-123 */
#define TSM_STEDR 6     /* core/stedr.py:14-86: EDR whose match also needs |t1 - t2| < delta */
#define TSM_STLCSS 7    /* core/stlcss.py:10-80: LCSS whose match also needs |t1 - t2| < delta */
#define TSM_CDDS 8      /* core/cdds.py:11-124: time the moving points stay within eps (needs t64) */
#define TSM_SAR 9       /* core/sar.py:10-286: time within eps on a common time grid, binned (needs t64) */

/* Error codes. */
#define TSM_OK 0
#define TSM_E_ARG (-1)     /* invalid argument */
#define TSM_E_HIP (-2)     /* HIP runtime error */
#define TSM_E_NOMEM (-3)   /* device allocation failed */
#define TSM_E_UNSUP (-4)   /* unsupported combination */

/* EDR/LCSS match conventions. */
#define TSM_EPS_SQRT 0    /* sqrt(dx^2+dy^2) < eps: traj_dist, core_cpu/stedr.pyx:11-12,36 */
#define TSM_EPS_SQUARED 1 /* dx^2+dy^2 < eps*eps: reference GPU, core/edr.py:22,42 */

/* Arithmetic of the DP. */
#define TSM_PREC_FP32 0 /* default: FP32 DP (EDR/LCSS match decisions are exact FP64 decisions) */
#define TSM_PREC_FP64 1 /* FP64 DP end to end (DTW, ERP, Frechet): the drift-check rerun */

/* A packed trajectory set resident on the device. */
typedef struct tsm_trajset {
    const float *xy32;      /* device [npts][2] FP32 copy of xy64, or of xy64 minus an origin
                               (tsm_xy_to_fp32 / tsm_xy_to_fp32_centred): the FP32 kernels use
                               coordinate differences only, so any origin gives the same results */
    const double *xy64;     /* device [npts][2] FP64 coordinates */
    const int64_t *offsets; /* device [n + 1], offsets[0] = 0 */
    int64_t n;              /* number of trajectories */
    int64_t npts;           /* number of points, offsets[n] */
    int64_t max_len;        /* longest trajectory (sizes scratch; must be >= every length) */
    double coord_max;       /* an upper bound on |x| and |y| of the FP32 copy (after its origin is
                               subtracted), or 0 if unknown.  EDR and
                               LCSS use it to run their match test on the FP32 copy with a proven
                               guard band (FP64 recheck inside the band); 0 keeps the test in FP64 */
    const double *t64;      /* device [npts] per-point timestamps for STEDR/STLCSS, or NULL */
} tsm_trajset;

/* Per-call parameters; tsm_params_default() fills the defaults. */
typedef struct tsm_params {
    double eps;     /* EDR/LCSS/STEDR/STLCSS/CDDS distance threshold (core/edr.py:55 `eps`) */
    int eps_mode;   /* TSM_EPS_SQRT (default) or TSM_EPS_SQUARED */
    double gx, gy;  /* ERP gap point g (core/erp.py:94 `gpoint`, default (0,0) erp.py:103) */
    int lanes;      /* lanes per pair: 0 = autotuned per column count, else 2, 4, 8 or 16 (threads_per_traj
                       hint; used where that many lanes hold the shorter trajectory's columns: at most 12
                       per lane with 2 lanes (16 for STEDR/STLCSS: 8), 24 otherwise) */
    int precision;  /* TSM_PREC_FP32 or TSM_PREC_FP64 */
    double delta;   /* STEDR/STLCSS time threshold (core/stedr.py:62 `delta`): match needs |t1 - t2| < delta */
    int64_t time_step;     /* SAR time slice (core/sar.py:247 `time_eps`), a positive integer; default 10 */
    int32_t target_length; /* SAR bins (core/sar.py:248 `target_length`), >= 1; default 10 */
    int32_t replicas;      /* SAR: the reference's threads_per_traj (core/sar.py:248, default 64).  Each of
                              its threads stores the same per-sample count when no binning happens
                              (core/sar.py:177-181, 236, 285), so that result is replicas * count *
                              time_step; 1 gives the single count.  Binned results do not depend on it. */
} tsm_params;

/* Library / device setup (reference: numba.cuda context, implicit). */
int tsm_init(int device);
const char *tsm_last_error(void);
int tsm_abi_version(void);
void tsm_params_default(tsm_params *p);

/* FP64 -> FP32 copy of a packed point buffer (2*npts values), stream-ordered. */
int tsm_xy_to_fp32(const double *xy64, float *xy32, int64_t npts, void *stream);

/* The same copy relative to an origin: xy32 = (float)(x - ox), (float)(y - oy), the
 * subtraction in FP64.  Mercator coordinates (~1e7 m, the dataset files of
 * preprocessing/pkl2h5py.py:16-27) lose ~0.5 m to a plain FP32 cast; relative to the
 * set's centre they keep FP32's precision of the set's extent.  The reference computes
 * in FP64 on whatever it is given (core/dtw.py:56-77), after its pairwise collate has
 * rounded coordinates to FP32 (utilities/method_helper.py:142-153); with the origin a
 * multiple of the coordinates' FP32 spacing such inputs stay exact.  coord_max of the
 * set must then bound |x - ox| and |y - oy|. */
int tsm_xy_to_fp32_centred(const double *xy64, float *xy32, int64_t npts, double ox, double oy, void *stream);

/*
 * Explicit pairs: pair k is (pa[k], pb[k]) of `set`, pa as t1 and pb as t2
 * (core/dtw.py:56 `t1, len1, t2, len2`).  pa, pb: device int32 [npairs].
 * out: device float64 [npairs].  counts: device int32 [npairs] or NULL
 * (EDR/LCSS only).  Replaces the kernel launch + copies of core/dtw.py:61-77,
 * core/edr.py:59-76, core/lcss.py:60-77, core/erp.py:98-120,
 * core/discret_frechet.py:57-74, core/hausdorff.py:101-123.
 */
int tsm_pairs(int measure, const tsm_trajset *set, const int32_t *pa, const int32_t *pb,
              int64_t npairs, const tsm_params *prm, double *out, int32_t *counts, void *stream);

/*
 * tsm_pairs for callers that cannot allocate device memory (code under
 * stream capture, serving loops and harnesses that forbid hipMalloc).  No
 * counterpart in the reference, which allocates per call (core/dtw.py:61-76).
 * Same arguments, kernels and results as tsm_pairs; it calls no hipMalloc,
 * hipFree or allocating routine.  The plan and the column-strip boundaries
 * of one call live in a fixed 32 MiB device array that each device holds for
 * the life of the process (a direct launch, tsm_tune_set_direct_pairs, and
 * the Hausdorff, CDDS and SAR kernels use none of it); a call needing more
 * (roughly 8 million pairs, fewer with trajectories longer than 256 points)
 * fails with TSM_E_ARG before launching anything.  Every call on a device shares
 * that array: calls that overlap on one device (other streams, other
 * threads) race and give undefined results.
 */
int tsm_pairs_noalloc(int measure, const tsm_trajset *set, const int32_t *pa, const int32_t *pb, int64_t npairs,
                      const tsm_params *prm, double *out, int32_t *counts, void *stream);

/*
 * Query x database grid: pair k = q * nd + d is (q0 + q, d0 + d) for
 * q < nq, d < nd, all indices into `set`.  Replaces the repeat/tile index
 * arrays and per-batch gathers of test_trajsimi_time_DQ.py:148-173.
 * out: device float64 [nq * nd], row-major.
 */
int tsm_pairs_grid(int measure, const tsm_trajset *set, int64_t q0, int64_t nq, int64_t d0, int64_t nd,
                   const tsm_params *prm, double *out, int32_t *counts, void *stream);

/*
 * Upper triangle of the all-pairs matrix over trajectories [0, n) of `set`:
 * pair k enumerates (i, j), i < j, row-major, like
 * itertools.combinations(range(n), 2) in test_clustering_time.py:219.
 * out: device float64 [n*(n-1)/2].
 */
int tsm_pairs_triu(int measure, const tsm_trajset *set, int64_t n, const tsm_params *prm, double *out,
                   int32_t *counts, void *stream);

/*
 * Host-memory convenience: uploads xy64 [npts][2] / offsets [n+1] (and the
 * timestamps t64 [npts] for STEDR/STLCSS, else NULL) and the pair lists,
 * runs tsm_pairs on the null stream, copies results back and synchronises.
 * This is the body of the reference's per-batch functions.
 */
int tsm_pairs_host(int measure, const double *xy64, const double *t64, const int64_t *offsets, int64_t n,
                   int64_t npts, const int32_t *pa, const int32_t *pb, int64_t npairs, const tsm_params *prm,
                   double *out, int32_t *counts);

/*
 * The reference's padded batches as they are: t1 [B][L1][dim] and t2
 * [B][L2][dim] (row-major float64, zero-padded; dim 2 = x, y or 3 = x, y, t)
 * with lengths len1[B], len2[B] in [1, L]; pair k is (t1[k], t2[k]).  Both
 * batches are copied to the device unchanged and compacted there (no host
 * repacking); timestamps are read for STEDR/STLCSS/CDDS/SAR (dim must be 3)
 * and ignored otherwise.  Non-finite values inside the lengths are an
 * error; padding is never read.  Runs on the null stream, copies the B
 * results (and counts) back and synchronises.  Replaces the body of the
 * reference's per-batch functions, which take exactly these arguments:
 * core/dtw.py:56-77 dtw(t1, len1, t2, len2), edr.py:55-76, lcss.py:56-77,
 * erp.py:94-120, discret_frechet.py:53-74, hausdorff.py:97-123,
 * stedr.py:65-86, stlcss.py:57-81, cdds.py:100-124, sar.py:247-286.
 */
int tsm_pairs_padded_host(int measure, const double *t1, const int64_t *len1, int64_t L1, const double *t2,
                          const int64_t *len2, int64_t L2, int64_t B, int dim, const tsm_params *prm, double *out,
                          int32_t *counts);

/*
 * Autotuner hook: lanes per pair (2, 4, 8 or 16; each lane then holds
 * ceil(ncol / lanes) register columns) for each column count 1..n-1 (entry 0
 * unused) of the DP kernels with FP32 (coord_bits 32) or FP64 (coord_bits
 * 64) coordinates; an entry whose lanes cannot hold its columns falls back
 * to the cost model, and NULL restores the cost model everywhere.  The
 * Python autotuner (trajsimimeasures_amd/autotune.py) measures and sets it.
 * No reference counterpart (the reference fixes 64 threads per pair,
 * config.py:45).
 */
int tsm_tune_set_lanes(int coord_bits, const uint8_t *lanes_of_ncol, int n);

/* Resident waves per CU of the persistent DP kernel (0 = default 8); the
 * autotuner measures and sets it.  No reference counterpart. */
int tsm_tune_set_waves_per_cu(int waves_per_cu);

/* Lanes per pair for pairs wider than one 16-lane strip (more than 384 columns
 * in the shorter trajectory): 64 (one wavefront: 12, 16, 20 or 24 columns per
 * lane, the fewest that hold the pair, then 24 per lane in strips of 1536; a
 * quarter of the steps of 16 lanes, so the longest pairs finish sooner) or 16 (strips of 384 columns: fewer
 * lane-steps per pair).  0 (default) picks 64 for launches of at most 2^20
 * pairs and 16 above.  No counterpart in the reference (it runs one thread per
 * column of one pair, core/dtw.py:33-45). */
int tsm_tune_set_wide_lanes(int lanes);

/* Direct DP launches: a call of at most max_pairs pairs, whose set needs no
 * column strips (max_len <= 64 * 24 columns with FP32 coordinates, 64 * 16 for
 * STEDR/STLCSS) and which passes no lanes hint, skips the planner and runs
 * one wavefront per pair (64 lanes of 4 to 24 columns, the fewest that hold
 * it): a small batch then lasts as long as its longest pair's chain of about
 * rows + columns / 4 steps instead of rows + 15 steps of wider lanes, and
 * pays no planner kernels.  -1 (default): three times the resident waves of
 * the persistent grid (3 x CUs x waves per CU, 6,144 on MI355X; the autotuner
 * measures the crossover); 0: never.  Results equal the planned launch's.  No counterpart in the reference, whose
 * batches of 512 pairs run 64 threads per pair (config.py:45,
 * method_helper.py:69-70). */
int tsm_tune_set_direct_pairs(int64_t max_pairs);

/*
 * hipGraph capture of one batch (planner and persistent DP kernel, or the
 * Hausdorff kernel) on a non-null stream: explicit pairs
 * (as tsm_pairs), a query x database grid (as tsm_pairs_grid: the per-batch
 * relaunches of test_trajsimi_time_DQ.py:148-173) or an all-pairs triangle
 * (as tsm_pairs_triu: test_clustering_time.py:219-254).  The capture call
 * runs the batch once eagerly (sizing the graph's private workspace and
 * producing results), then records the same launches; each tsm_graph_launch
 * replays them on the same device buffers.  For the reference's repeated
 * fixed-size batches (effi_batch_size_gpu = 512, config.py:43) and serving
 * loops.  The graph holds no memsets.
 *
 * What a replay reads afresh and what the capture froze:
 *   - read afresh: the pair index arrays, the offsets (lengths may change if
 *     npts stays the same and no trajectory exceeds the captured max_len),
 *     xy64, xy32 and t64;
 *   - frozen: the pair count or grid / triangle ranges, every tsm_params
 *     field, the set's pointers, and the EDR/LCSS match path chosen from the
 *     set's coord_max: the FP32 test's guard band is derived from that bound.
 * So refilled coordinates must keep xy32 = (float)(xy64 - origin) for the
 * origin the FP32 copy was made with (tsm_xy_to_fp32_centred), with every
 * |xy32| within the captured coord_max; beyond it EDR/LCSS counts are no
 * longer guaranteed exact.  Refilling xy64 alone leaves the FP32 measures on
 * the old points.
 */
typedef struct tsm_graph tsm_graph;
int tsm_graph_capture_pairs(int measure, const tsm_trajset *set, const int32_t *pa, const int32_t *pb,
                            int64_t npairs, const tsm_params *prm, double *out, int32_t *counts, void *stream,
                            tsm_graph **graph);
int tsm_graph_capture_grid(int measure, const tsm_trajset *set, int64_t q0, int64_t nq, int64_t d0, int64_t nd,
                           const tsm_params *prm, double *out, int32_t *counts, void *stream, tsm_graph **graph);
int tsm_graph_capture_triu(int measure, const tsm_trajset *set, int64_t n, const tsm_params *prm, double *out,
                           int32_t *counts, void *stream, tsm_graph **graph);
int tsm_graph_launch(tsm_graph *graph, void *stream);
int tsm_graph_destroy(tsm_graph *graph);

/* Kernel timing of the DP / Hausdorff / CDDS / SAR kernels (HIP events
 * stamped by the kernel's own dispatch, on its stream), for roofline
 * reporting.  tsm_profile_enable(n): time every n-th launch from now on
 * (n = 1: all; n = 0: off; a timed dispatch costs a few microseconds, so a
 * stride keeps repeated identical steps near their untimed speed).
 * tsm_profile_read: total milliseconds and number of the timed launches
 * since the last read or enable (waits for them).  No reference
 * counterpart. */
int tsm_profile_enable(int every);
int tsm_profile_read(double *total_ms, int64_t *launches);

#ifdef __cplusplus
}
#endif
#endif /* TSM_H */
