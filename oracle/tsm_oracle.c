/*
 * tsm_oracle.c -- FP64 CPU oracle for the six trajectory-similarity measures.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: the product path
 * (trajsimimeasures_amd, libtsm.so) never links, loads or calls it.  Only
 * tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use it.
 *
 * Each measure is a textbook full-matrix (n1+1) x (n2+1) dynamic program in
 * float64, the shape of the traj_dist Cython routines the reference uses as
 * its CPU path (utilities/method_helper.py:45-50, traj_dist==1.15), with the
 * boundary conditions and quirks of the reference's GPU kernels in core/{dtw,edr,...}.py,
 * which this engine is a drop-in for:
 *
 *   DTW       core/dtw.py:27-53             inf boundary, D = d + min3
 *   EDR       core/edr.py:29-52             ZERO boundary, min(diag+sub, up+1, left+1), / max(n1,n2)
 *   LCSS      core/lcss.py:31-53            zero boundary, match -> diag+1 else max(up,left), 1 - c/min(n1,n2)
 *   ERP       core/erp.py:29-91             constant-sum boundary D[i][0]=Sa, D[0][j]=Sb
 *   Frechet   core/discret_frechet.py:26-50 inf boundary, D = max(d, min3)
 *   Hausdorff core/hausdorff.py:12-69,121   point-to-segment, squared distances, sqrt at the end
 *   STEDR     core/stedr.py:29-59           EDR whose match also needs |t1 - t2| < delta
 *   STLCSS    core/stlcss.py:27-54          LCSS whose match also needs |t1 - t2| < delta
 *             (CPU forms: core_cpu/stedr.pyx:15-42, core_cpu/stlcss.pyx:15-39)
 *   CDDS      core/cdds.py:11-94            sum over all segment pairs of the time the two
 *             moving points are within eps (CPU form: core_cpu/cdds.pyx)
 *   SAR       core/sar.py:10-212            time-aligned, binned comparison of the overlapping
 *             stretch (CPU form: core_cpu/sar.pyx)
 *   k-medoids test_clustering_time.py:353-367   FasterPAM over the all-pairs matrix
 *
 * Distances are sqrt(dx*dx + dy*dy) evaluated without FMA contraction
 * (core/utils.py:8-17); build with -ffp-contract=off.
 *
 * The EDR/LCSS match test is supplied as a squared-distance threshold `thr`:
 * a cell matches iff dx*dx + dy*dy < thr.  tsmo_eps_threshold() turns an
 * epsilon into that threshold for both conventions:
 *   mode 0 ("sqrt", traj_dist / core_cpu/stedr.pyx:11-12,36):  sqrt(d2) < eps
 *   mode 1 ("squared", reference GPU, core/edr.py:22,42):       d2 < eps*eps
 * For mode 0 the threshold is the least double t with sqrt(t) >= eps, so
 * d2 < t  <=>  sqrt(d2) < eps for every double d2 (sqrt is correctly rounded
 * and monotone).
 *
 * The batched entry tsmo_pairs() runs one pair per task with OpenMP over
 * contiguous slices, like the reference CPU driver's slicing
 * (utilities/tool_funcs.py:195-208, test_trajsimi_time.py:280-316).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

enum { TSMO_DTW = 0, TSMO_EDR = 1, TSMO_LCSS = 2, TSMO_ERP = 3, TSMO_FRECHET = 4, TSMO_HAUSDORFF = 5,
       TSMO_STEDR = 6, TSMO_STLCSS = 7, TSMO_CDDS = 8, TSMO_SAR = 9 };

static inline double sqd(const double *p, const double *q) {
    double dx = p[0] - q[0];
    double dy = p[1] - q[1];
    return dx * dx + dy * dy;
}
static inline double eucl(const double *p, const double *q) { return sqrt(sqd(p, q)); }
static inline double min2(double a, double b) { return a < b ? a : b; }
static inline double max2(double a, double b) { return a > b ? a : b; }
static inline double min3(double a, double b, double c) { return min2(min2(a, b), c); }

/* Least double t with sqrt(t) >= eps (mode 0), or eps*eps (mode 1). */
double tsmo_eps_threshold(double eps, int mode) {
    if (mode == 1) return eps * eps;
    if (!(eps > 0.0)) return 0.0; /* sqrt(d2) < eps impossible for eps <= 0 */
    if (isinf(eps)) return INFINITY;
    double t = eps * eps;
    while (sqrt(t) >= eps) t = nextafter(t, 0.0);
    while (sqrt(t) < eps) t = nextafter(t, INFINITY);
    return t;
}

/* core/dtw.py:27-53 -- D[0][0]=0, D[i][0]=D[0][j]=inf, D=d+min3. */
double tsmo_dtw(const double *a, int64_t n1, const double *b, int64_t n2) {
    const int64_t w = n2 + 1;
    double *D = (double *)malloc(sizeof(double) * (size_t)(n1 + 1) * (size_t)w);
    for (int64_t j = 0; j <= n2; ++j) D[j] = INFINITY;
    D[0] = 0.0;
    for (int64_t i = 1; i <= n1; ++i) {
        D[i * w] = INFINITY;
        for (int64_t j = 1; j <= n2; ++j) {
            double d = eucl(a + 2 * (i - 1), b + 2 * (j - 1));
            D[i * w + j] = d + min3(D[(i - 1) * w + j - 1], D[(i - 1) * w + j], D[i * w + j - 1]);
        }
    }
    double r = D[n1 * w + n2];
    free(D);
    return r;
}

/* core/discret_frechet.py:26-50 -- same boundary as DTW, D = max(d, min3). */
double tsmo_frechet(const double *a, int64_t n1, const double *b, int64_t n2) {
    const int64_t w = n2 + 1;
    double *D = (double *)malloc(sizeof(double) * (size_t)(n1 + 1) * (size_t)w);
    for (int64_t j = 0; j <= n2; ++j) D[j] = INFINITY;
    D[0] = 0.0;
    for (int64_t i = 1; i <= n1; ++i) {
        D[i * w] = INFINITY;
        for (int64_t j = 1; j <= n2; ++j) {
            double d = eucl(a + 2 * (i - 1), b + 2 * (j - 1));
            D[i * w + j] = max2(d, min3(D[(i - 1) * w + j - 1], D[(i - 1) * w + j], D[i * w + j - 1]));
        }
    }
    double r = D[n1 * w + n2];
    free(D);
    return r;
}

/* core/edr.py:29-52 -- zero boundary; returns the integer count C[n1][n2]. */
int64_t tsmo_edr_count(const double *a, int64_t n1, const double *b, int64_t n2, double thr) {
    const int64_t w = n2 + 1;
    int64_t *C = (int64_t *)calloc((size_t)(n1 + 1) * (size_t)w, sizeof(int64_t));
    for (int64_t i = 1; i <= n1; ++i) {
        for (int64_t j = 1; j <= n2; ++j) {
            int64_t sub = sqd(a + 2 * (i - 1), b + 2 * (j - 1)) < thr ? 0 : 1;
            int64_t v = C[(i - 1) * w + j - 1] + sub;
            int64_t up = C[(i - 1) * w + j] + 1;
            int64_t left = C[i * w + j - 1] + 1;
            if (up < v) v = up;
            if (left < v) v = left;
            C[i * w + j] = v;
        }
    }
    int64_t r = C[n1 * w + n2];
    free(C);
    return r;
}

/* core/lcss.py:31-53 -- zero boundary; returns the integer count C[n1][n2]. */
int64_t tsmo_lcss_count(const double *a, int64_t n1, const double *b, int64_t n2, double thr) {
    const int64_t w = n2 + 1;
    int64_t *C = (int64_t *)calloc((size_t)(n1 + 1) * (size_t)w, sizeof(int64_t));
    for (int64_t i = 1; i <= n1; ++i) {
        for (int64_t j = 1; j <= n2; ++j) {
            int64_t v;
            if (sqd(a + 2 * (i - 1), b + 2 * (j - 1)) < thr) {
                v = C[(i - 1) * w + j - 1] + 1;
            } else {
                int64_t up = C[(i - 1) * w + j], left = C[i * w + j - 1];
                v = up > left ? up : left;
            }
            C[i * w + j] = v;
        }
    }
    int64_t r = C[n1 * w + n2];
    free(C);
    return r;
}

/* Spatio-temporal match: (dx*dx + dy*dy) < thr and |ta - tb| < delta
 * (core/stedr.py:44-46, core/stlcss.py:41-42; fabs as core_cpu/stedr.pyx:37). */
static inline int st_match(const double *p, double tp, const double *q, double tq, double thr, double delta) {
    return sqd(p, q) < thr && fabs(tp - tq) < delta;
}

/* core/stedr.py:29-59 -- EDR on spatio-temporal matches; returns the count. */
int64_t tsmo_stedr_count(const double *a, const double *ta, int64_t n1, const double *b, const double *tb,
                         int64_t n2, double thr, double delta) {
    const int64_t w = n2 + 1;
    int64_t *C = (int64_t *)calloc((size_t)(n1 + 1) * (size_t)w, sizeof(int64_t));
    for (int64_t i = 1; i <= n1; ++i) {
        for (int64_t j = 1; j <= n2; ++j) {
            int64_t sub = st_match(a + 2 * (i - 1), ta[i - 1], b + 2 * (j - 1), tb[j - 1], thr, delta) ? 0 : 1;
            int64_t v = C[(i - 1) * w + j - 1] + sub;
            int64_t up = C[(i - 1) * w + j] + 1;
            int64_t left = C[i * w + j - 1] + 1;
            if (up < v) v = up;
            if (left < v) v = left;
            C[i * w + j] = v;
        }
    }
    int64_t r = C[n1 * w + n2];
    free(C);
    return r;
}

/* core/stlcss.py:27-54 -- LCSS on spatio-temporal matches; returns the count. */
int64_t tsmo_stlcss_count(const double *a, const double *ta, int64_t n1, const double *b, const double *tb,
                          int64_t n2, double thr, double delta) {
    const int64_t w = n2 + 1;
    int64_t *C = (int64_t *)calloc((size_t)(n1 + 1) * (size_t)w, sizeof(int64_t));
    for (int64_t i = 1; i <= n1; ++i) {
        for (int64_t j = 1; j <= n2; ++j) {
            int64_t v;
            if (st_match(a + 2 * (i - 1), ta[i - 1], b + 2 * (j - 1), tb[j - 1], thr, delta)) {
                v = C[(i - 1) * w + j - 1] + 1;
            } else {
                int64_t up = C[(i - 1) * w + j], left = C[i * w + j - 1];
                v = up > left ? up : left;
            }
            C[i * w + j] = v;
        }
    }
    int64_t r = C[n1 * w + n2];
    free(C);
    return r;
}

double tsmo_edr(const double *a, int64_t n1, const double *b, int64_t n2, double thr) {
    return (double)tsmo_edr_count(a, n1, b, n2, thr) / (double)(n1 > n2 ? n1 : n2); /* core/edr.py:52 */
}

double tsmo_lcss(const double *a, int64_t n1, const double *b, int64_t n2, double thr) {
    return 1.0 - (double)tsmo_lcss_count(a, n1, b, n2, thr) / (double)(n1 < n2 ? n1 : n2); /* core/lcss.py:53 */
}

/* core/erp.py:29-91 -- gap distances to g, constant boundary sums Sa, Sb
 * (erp.py:44-49 serial sums, erp.py:62-78 boundary use). */
double tsmo_erp(const double *a, int64_t n1, const double *b, int64_t n2, const double *g) {
    const int64_t w = n2 + 1;
    double *ga = (double *)malloc(sizeof(double) * (size_t)n1);
    double *gb = (double *)malloc(sizeof(double) * (size_t)n2);
    double sa = 0.0, sb = 0.0;
    for (int64_t i = 0; i < n1; ++i) { ga[i] = eucl(g, a + 2 * i); }
    for (int64_t j = 0; j < n2; ++j) { gb[j] = eucl(g, b + 2 * j); }
    for (int64_t i = 0; i < n1; ++i) sa += ga[i];
    for (int64_t j = 0; j < n2; ++j) sb += gb[j];
    double *D = (double *)malloc(sizeof(double) * (size_t)(n1 + 1) * (size_t)w);
    D[0] = 0.0;
    for (int64_t j = 1; j <= n2; ++j) D[j] = sb;
    for (int64_t i = 1; i <= n1; ++i) {
        D[i * w] = sa;
        for (int64_t j = 1; j <= n2; ++j) {
            double left = D[i * w + j - 1] + gb[j - 1];
            double up = D[(i - 1) * w + j] + ga[i - 1];
            double diag = D[(i - 1) * w + j - 1] + eucl(a + 2 * (i - 1), b + 2 * (j - 1));
            D[i * w + j] = min3(left, up, diag);
        }
    }
    double r = D[n1 * w + n2];
    free(D);
    free(ga);
    free(gb);
    return r;
}

/* core/hausdorff.py:12-45 -- squared point-to-segment distance. */
static double p2s_sq(const double *p, const double *s1, const double *s2, double dps1, double dps2) {
    double segl = sqd(s1, s2);
    if (s1[0] == s2[0] && s1[1] == s2[1]) return dps1;
    double xd = s2[0] - s1[0];
    double yd = s2[1] - s1[1];
    double u1 = ((p[0] - s1[0]) * xd) + ((p[1] - s1[1]) * yd);
    double u = u1 / segl;
    if (u < 0.00001 || u > 1.0) return min2(dps1, dps2);
    double ix = s1[0] + u * xd;
    double iy = s1[1] + u * yd;
    double dx = p[0] - ix, dy = p[1] - iy;
    return dx * dx + dy * dy;
}

/* core/hausdorff.py:48-69 -- squared point-to-polyline distance; 0 when the
 * polyline has a single point (no segment), as the reference returns. */
static double p2t_sq(const double *p, const double *t, int64_t n) {
    double dist = 0.0, dps1 = 0.0, dps2 = 0.0;
    for (int64_t i = 0; i + 1 < n; ++i) {
        if (i == 0) {
            dps1 = sqd(p, t);
            dps2 = sqd(p, t + 2);
        } else {
            dps1 = dps2;
            dps2 = sqd(p, t + 2 * (i + 1));
        }
        double ps = p2s_sq(p, t + 2 * i, t + 2 * (i + 1), dps1, dps2);
        if (i == 0 || ps < dist) dist = ps;
    }
    return dist;
}

/* core/hausdorff.py:97-123 -- sqrt of the max over both directions, padding
 * excluded (the reference reduces over uninitialised padding, :114,121). */
double tsmo_hausdorff(const double *a, int64_t n1, const double *b, int64_t n2) {
    double m = 0.0;
    for (int64_t i = 0; i < n1; ++i) m = max2(m, p2t_sq(a + 2 * i, b, n2));
    for (int64_t j = 0; j < n2; ++j) m = max2(m, p2t_sq(b + 2 * j, a, n1));
    return sqrt(m);
}

/* One pair of the batched oracle; count_out receives the EDR/LCSS count. */
static double one_pair(int measure, const double *a, int64_t n1, const double *b, int64_t n2,
                       double thr, const double *g, int64_t *count_out) {
    switch (measure) {
    case TSMO_DTW: return tsmo_dtw(a, n1, b, n2);
    case TSMO_EDR: {
        int64_t c = tsmo_edr_count(a, n1, b, n2, thr);
        *count_out = c;
        return (double)c / (double)(n1 > n2 ? n1 : n2);
    }
    case TSMO_LCSS: {
        int64_t c = tsmo_lcss_count(a, n1, b, n2, thr);
        *count_out = c;
        return 1.0 - (double)c / (double)(n1 < n2 ? n1 : n2);
    }
    case TSMO_ERP: return tsmo_erp(a, n1, b, n2, g);
    case TSMO_FRECHET: return tsmo_frechet(a, n1, b, n2);
    case TSMO_HAUSDORFF: return tsmo_hausdorff(a, n1, b, n2);
    default: return NAN;
    }
}

/*
 * Batched oracle over a packed trajectory set: xy is float64 [npts, 2],
 * offsets int64 [ntraj + 1]; pair k is (pa[k], pb[k]) with pa as t1.
 * out[k] receives the measure value, counts[k] (if non-NULL) the EDR/LCSS
 * count.  nthreads <= 0 uses the OpenMP default.  Returns 0, or -1 on a bad
 * measure id.
 */
int tsmo_pairs(int measure, const double *xy, const int64_t *offsets, const int64_t *pa,
               const int64_t *pb, int64_t npairs, double thr, const double *g, double *out,
               int64_t *counts, int nthreads) {
    if (measure < 0 || measure > 5) return -1;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    int64_t k;
#pragma omp parallel for schedule(static)
    for (k = 0; k < npairs; ++k) {
        int64_t ia = pa[k], ib = pb[k];
        int64_t c = -1;
        const double *a = xy + 2 * offsets[ia];
        const double *b = xy + 2 * offsets[ib];
        out[k] = one_pair(measure, a, offsets[ia + 1] - offsets[ia], b, offsets[ib + 1] - offsets[ib], thr, g, &c);
        if (counts) counts[k] = c;
    }
    return 0;
}

/* Batched STEDR (6) / STLCSS (7) over pairs of a set with per-point times
 * t[npts]; values normalised as core/stedr.py:57 and core/stlcss.py:52. */
int tsmo_pairs_st(int measure, const double *xy, const double *t, const int64_t *offsets, const int64_t *pa,
                  const int64_t *pb, int64_t npairs, double thr, double delta, double *out, int64_t *counts,
                  int nthreads) {
    if (measure != TSMO_STEDR && measure != TSMO_STLCSS) return -1;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    int64_t k;
#pragma omp parallel for schedule(static)
    for (k = 0; k < npairs; ++k) {
        const int64_t ia = pa[k], ib = pb[k];
        const int64_t oa = offsets[ia], ob = offsets[ib];
        const int64_t n1 = offsets[ia + 1] - oa, n2 = offsets[ib + 1] - ob;
        int64_t c;
        if (measure == TSMO_STEDR) {
            c = tsmo_stedr_count(xy + 2 * oa, t + oa, n1, xy + 2 * ob, t + ob, n2, thr, delta);
            out[k] = (double)c / (double)(n1 > n2 ? n1 : n2);
        } else {
            c = tsmo_stlcss_count(xy + 2 * oa, t + oa, n1, xy + 2 * ob, t + ob, n2, thr, delta);
            out[k] = 1.0 - (double)c / (double)(n1 < n2 ? n1 : n2);
        }
        if (counts) counts[k] = c;
    }
    return 0;
}

/* core/cdds.py:11-73 (cdd) -- time during which the points moving linearly
 * along segment e1 = (p, q) and e2 = (r, s) (each point x, y, t) are within
 * eps of each other.  Positions are x(t) = xk t + xb (core/cdds.py:11-18);
 * the squared distance a t^2 + b t + c is solved against eps^2.  A segment
 * with no duration contributes 0 (core_cpu/cdds.pyx:26; the reference GPU
 * kernel would divide by zero there). */
static void slope_intercept(const double *p, const double *q, double *xk, double *xb, double *yk, double *yb) {
    const double dt = q[2] - p[2];
    *yk = (q[1] - p[1]) / dt;
    *yb = (p[1] * q[2] - q[1] * p[2]) / dt;
    *xk = (q[0] - p[0]) / dt;
    *xb = (p[0] * q[2] - q[0] * p[2]) / dt;
}

double tsmo_cdd(const double *e1p1, const double *e1p2, const double *e2p1, const double *e2p2, double eps) {
    if (e1p2[2] <= e2p1[2] || e2p2[2] <= e1p1[2] || e1p1[2] == e1p2[2] || e2p1[2] == e2p2[2]) return 0.0;
    const double se = eps * eps;
    double axk, axb, ayk, ayb, bxk, bxb, byk, byb;
    slope_intercept(e1p1, e1p2, &axk, &axb, &ayk, &ayb);
    slope_intercept(e2p1, e2p2, &bxk, &bxb, &byk, &byb);
    const double dxk = axk - bxk, dyk = ayk - byk, dxb = axb - bxb, dyb = ayb - byb;
    const double a = dxk * dxk + dyk * dyk;
    const double b = 2.0 * (dxk * dxb + dyk * dyb);
    double c = dxb * dxb + dyb * dyb;
    const double t_start = max2(e1p1[2], e2p1[2]);
    const double t_end = min2(e1p2[2], e2p2[2]);
    if (a == 0.0 && b == 0.0 && c == 0.0) return t_end - t_start;
    if (a == 0.0 && b == 0.0) return c <= se ? t_end - t_start : 0.0;
    if (a == 0.0) {
        const double t1 = -1.0 * c / b, t2 = (se - c) / b;
        if (t1 < t2) {
            const double lo = max2(t1, t_start), hi = min2(t2, t_end);
            return hi > lo ? hi - lo : 0.0;
        }
        if (t2 < t1) {
            const double lo = max2(t2, t_start), hi = min2(t1, t_end);
            return hi > lo ? hi - lo : 0.0;
        }
        return 0.0;
    }
    c -= se;
    const double theta = b * b - 4.0 * a * c;
    if (theta > 0.0) {
        const double ts = sqrt(theta);
        const double t1 = (-1.0 * ts - b) / 2.0 / a, t2 = (ts - b) / 2.0 / a;
        const double lo = max2(t1, t_start), hi = min2(t2, t_end);
        return hi > lo ? hi - lo : 0.0;
    }
    return 0.0;
}

/* core/cdds.py:76-94,124 -- for each segment i of a: the sum over the
 * segments j of b in order; then the sum over i in order. */
double tsmo_cdds(const double *a, const double *ta, int64_t n1, const double *b, const double *tb, int64_t n2,
                 double eps) {
    double total = 0.0;
    for (int64_t i = 0; i + 1 < n1; ++i) {
        const double p[3] = {a[2 * i], a[2 * i + 1], ta[i]}, q[3] = {a[2 * i + 2], a[2 * i + 3], ta[i + 1]};
        double d = 0.0;
        for (int64_t j = 0; j + 1 < n2; ++j) {
            const double r[3] = {b[2 * j], b[2 * j + 1], tb[j]}, s[3] = {b[2 * j + 2], b[2 * j + 3], tb[j + 1]};
            d += tsmo_cdd(p, q, r, s, eps);
        }
        total += d;
    }
    return total;
}

int tsmo_pairs_cdds(const double *xy, const double *t, const int64_t *offsets, const int64_t *pa, const int64_t *pb,
                    int64_t npairs, double eps, double *out, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    int64_t k;
#pragma omp parallel for schedule(static)
    for (k = 0; k < npairs; ++k) {
        const int64_t oa = offsets[pa[k]], ob = offsets[pb[k]];
        out[k] = tsmo_cdds(xy + 2 * oa, t + oa, offsets[pa[k] + 1] - oa, xy + 2 * ob, t + ob,
                           offsets[pb[k] + 1] - ob, eps);
    }
    return 0;
}

/* ---- SAR (core/sar.py) ----------------------------------------------------
 * The reference GPU form, step by step:
 *  1. sar_preprocess (core/sar.py:62-103): a merge over both trajectories'
 *     segments in order of end time finds the first and last segments that
 *     overlap the other trajectory in time (open intervals, zero-duration
 *     segments excluded); no overlap gives 0.
 *  2. align_to_timestamps (core/sar.py:31-59, 10-28): the overlapping stretch
 *     of each trajectory is sampled at every multiple of time_eps, segment by
 *     segment (grid indices ceil(t_k/time_eps) .. ceil(t_k+1/time_eps) - 1,
 *     positions interpolated as x = xk*ts + xb), plus the stretch's last point
 *     when its time is a multiple of time_eps.  The samples are stored in
 *     FP32 (interpolated_points_g is float32, core/sar.py:271).
 *  3. Trimming (core/sar.py:118-150): the later end loses
 *     (difference // time_eps) samples, then the earlier start does; the time
 *     differences are FP32 (the samples' type) and // is Python's floor
 *     division.
 *  4. sar_compute (core/sar.py:163-212): with sz samples left and
 *     target_length M >= sz, count the sample pairs within distance_eps (FP32
 *     differences, sqrt <= distance_eps) times time_eps; every one of the
 *     reference's threads_per_traj threads computes and stores that same
 *     count and the host sums them (core/sar.py:236,285), so the result is
 *     threads_per_traj * count * time_eps.  With M < sz: M bins, bin b
 *     covering samples [b*div + min(b, mod), (b+1)*div + min(b+1, mod) + 1)
 *     (div, mod = sz // M, sz % M; note the "+ 1", core/sar.py:195), FP64
 *     means in sample order; a bin whose means are within distance_eps adds
 *     its sample count times time_eps.
 * Two reads of the reference fall outside the samples it holds, and are
 * defined here: the last bin's "+ 1" sample lies past the trimmed stretch
 * (uninitialised device memory for at least one side), so the last bin ends
 * at sz; and when the two trimmed stretches differ in length (which needs
 * timestamps past 2^24 or decreasing ones) the shorter length is used.
 * A pair with no overlap returns 0 (the reference leaves its slot
 * unwritten).  A stretch of more than 2^24 samples gives NaN. */
typedef struct { float x, y, t; } sar_pt;

#define SAR_MAX_SAMPLES (1ll << 24)

/* float -> int64 as the conversion of the reference's math.ceil / int():
 * truncation, saturating, NaN -> 0 */
static int64_t sar_to_i64(double v) {
    if (v != v) return 0;
    if (v >= 9223372036854775807.0) return INT64_MAX;
    if (v <= -9223372036854775808.0) return INT64_MIN;
    return (int64_t)v;
}

/* Python's float floor division a // b (the algorithm of CPython's
 * float_floor_div, which Numba follows): fmod first, so that an exact
 * multiple never rounds down. */
static double sar_floordiv(double a, double b) {
    double mod = fmod(a, b);
    double div = (a - mod) / b;
    if (mod != 0.0 && ((b < 0.0) != (mod < 0.0))) div -= 1.0;
    double fl;
    if (div != 0.0) {
        fl = floor(div);
        if (div - fl > 0.5) fl += 1.0;
    } else {
        fl = copysign(0.0, a / b);
    }
    return fl;
}

/* align_to_timestamps over points [0, n) of (xy, t); writes up to cap
 * samples to out (NULL: count only); returns the sample count. */
static int64_t sar_align(const double *xy, const double *t, int64_t n, int64_t ts, sar_pt *out, int64_t cap) {
    const double tsd = (double)ts;
    int64_t c = 0;
    for (int64_t i = 0; i + 1 < n; ++i) {
        const double t1 = t[i], t2 = t[i + 1];
        const int64_t s = sar_to_i64(ceil(t1 / tsd)), e = sar_to_i64(ceil(t2 / tsd));
        if (s >= e) continue;
        if (e - s > cap - c) return cap + 1;
        const double x1 = xy[2 * i], y1 = xy[2 * i + 1], x2 = xy[2 * i + 2], y2 = xy[2 * i + 3];
        const double yk = (y2 - y1) / (t2 - t1), yb = (y1 * t2 - y2 * t1) / (t2 - t1);
        const double xk = (x2 - x1) / (t2 - t1), xb = (x1 * t2 - x2 * t1) / (t2 - t1);
        for (int64_t g = s; g < e; ++g, ++c) {
            if (!out) continue;
            const int64_t tt = g * ts;
            out[c].x = (float)(xk * (double)tt + xb);
            out[c].y = (float)(yk * (double)tt + yb);
            out[c].t = (float)tt;
        }
    }
    if (fmod(t[n - 1], tsd) == 0.0) {
        if (c + 1 > cap) return cap + 1;
        if (out) {
            out[c].x = (float)xy[2 * (n - 1)];
            out[c].y = (float)xy[2 * (n - 1) + 1];
            out[c].t = (float)t[n - 1];
        }
        ++c;
    }
    return c;
}

static int64_t min_i64(int64_t a, int64_t b) { return a < b ? a : b; }

double tsmo_sar(const double *a, const double *ta, int64_t n1, const double *b, const double *tb, int64_t n2,
                double deps, int64_t ts, int64_t M, int64_t R) {
    /* 1. overlap merge (core/sar.py:79-100) */
    int64_t i = 0, j = 0, is = -1, ie = -1, js = -1, je = -1;
    while (i < n1 - 1 && j < n2 - 1) {
        const double a0 = ta[i], a1 = ta[i + 1], b0 = tb[j], b1 = tb[j + 1];
        if (!(a1 <= b0 || b1 <= a0 || a0 == a1 || b0 == b1)) {
            if (is < 0) is = i;
            if (js < 0) js = j;
            ie = i + 1;
            je = j + 1;
        }
        if (a1 < b1) ++i;
        else if (b1 < a1) ++j;
        else { ++i; ++j; }
    }
    if (is < 0) return 0.0;
    /* 2. samples of both stretches */
    const int64_t c1 = sar_align(a + 2 * is, ta + is, ie - is + 1, ts, NULL, SAR_MAX_SAMPLES);
    const int64_t c2 = sar_align(b + 2 * js, tb + js, je - js + 1, ts, NULL, SAR_MAX_SAMPLES);
    if (c1 > SAR_MAX_SAMPLES || c2 > SAR_MAX_SAMPLES) return NAN;
    if (c1 == 0 || c2 == 0) return 0.0;
    sar_pt *A = (sar_pt *)malloc(sizeof(sar_pt) * (size_t)(c1 + c2));
    if (!A) return NAN;
    sar_pt *B = A + c1;
    sar_align(a + 2 * is, ta + is, ie - is + 1, ts, A, c1);
    sar_align(b + 2 * js, tb + js, je - js + 1, ts, B, c2);
    /* 3. trimming (core/sar.py:118-150) */
    const float as = A[0].t, ae = A[c1 - 1].t, bs = B[0].t, be = B[c2 - 1].t;
    const double tsd = (double)ts;
    int64_t s1 = 0, e1 = c1, s2 = 0, e2 = c2, z1 = c1, z2 = c2;
    if (ae == be) {
    } else if (ae < be) {
        const int64_t k = sar_to_i64(sar_floordiv((double)(float)(be - ae), tsd));
        e2 -= min_i64(k, z2);
        z2 = e2 - s2;
    } else if (ae > be) {
        const int64_t k = sar_to_i64(sar_floordiv((double)(float)(ae - be), tsd));
        e1 -= min_i64(k, z1);
        z1 = e1 - s1;
    }
    double result = 0.0;
    if (z1 != 0 && z2 != 0) {
        if (as == bs) {
        } else if (as < bs) {
            const int64_t k = sar_to_i64(sar_floordiv((double)(float)(bs - as), tsd));
            s1 += min_i64(k, c1);
            z1 = e1 - s1;
        } else if (as > bs) {
            const int64_t k = sar_to_i64(sar_floordiv((double)(float)(as - bs), tsd));
            s2 += min_i64(k, c2);
            z2 = e2 - s2;
        }
        const int64_t sz = min_i64(z1, z2);
        /* 4. comparison (core/sar.py:163-212) */
        if (sz > 0 && M >= sz) {
            int64_t cnt = 0;
            for (int64_t m = 0; m < sz; ++m) {
                const float dx = B[s2 + m].x - A[s1 + m].x, dy = B[s2 + m].y - A[s1 + m].y;
                const float d2 = dx * dx + dy * dy;
                if ((double)(float)sqrt((double)d2) <= deps) ++cnt;
            }
            result = (double)(R * cnt * ts);
        } else if (sz > 0) {
            const int64_t dv = sz / M, md = sz % M;
            int64_t total = 0;
            for (int64_t bi = 0; bi < M; ++bi) {
                const int64_t st = bi * dv + min_i64(bi, md);
                int64_t en = (bi + 1) * dv + min_i64(bi + 1, md) + 1;
                if (en > sz) en = sz;
                double ax = 0.0, ay = 0.0, bx = 0.0, by = 0.0;
                for (int64_t k = st; k < en; ++k) {
                    ax += A[s1 + k].x;
                    ay += A[s1 + k].y;
                    bx += B[s2 + k].x;
                    by += B[s2 + k].y;
                }
                const double cn = (double)(en - st);
                ax /= cn;
                ay /= cn;
                bx /= cn;
                by /= cn;
                if (sqrt((bx - ax) * (bx - ax) + (by - ay) * (by - ay)) <= deps) total += (en - st) * ts;
            }
            result = (double)total;
        }
    }
    free(A);
    return result;
}

int tsmo_pairs_sar(const double *xy, const double *t, const int64_t *offsets, const int64_t *pa, const int64_t *pb,
                   int64_t npairs, double deps, int64_t ts, int64_t M, int64_t R, double *out, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    int64_t k;
#pragma omp parallel for schedule(static)
    for (k = 0; k < npairs; ++k) {
        const int64_t oa = offsets[pa[k]], ob = offsets[pb[k]];
        out[k] = tsmo_sar(xy + 2 * oa, t + oa, offsets[pa[k] + 1] - oa, xy + 2 * ob, t + ob,
                          offsets[pb[k] + 1] - ob, deps, ts, M, R);
    }
    return 0;
}

/* ---- k-medoids: FasterPAM -------------------------------------------------
 * The consumer of the all-pairs matrix, test_clustering_time.py:353-367:
 * kmedoids.fasterpam(dists, Config.clustering_num_centroids) from the
 * kmedoids package (kmedoids==0.4.3, requirements377.txt:6; a Rust
 * extension, absent here).  Restated from the published algorithm (E.
 * Schubert, P. J. Rousseeuw, "Fast and eager k-medoids clustering: O(k)
 * runtime improvement of the PAM, CLARA, and CLARANS algorithms",
 * Information Systems 101 (2021), FasterPAM):
 *  - every point keeps its nearest medoid (an index into the medoid list and
 *    the distance) and its second nearest; a point that is itself a medoid
 *    is assigned to itself;
 *  - removal loss of medoid m: the sum over its points of (second - nearest);
 *  - candidates x = 0 .. n-1 in order (medoids skipped): one pass over the
 *    points gives, for every medoid m, the loss change of swapping m out and
 *    x in: points for which x is nearer than their nearest add the shared
 *    d(x,o) - dn and take back their removal term (dn - ds) from m = n(o);
 *    points for which x lies between nearest and second change the removal
 *    term of n(o) to d(x,o) - ds.  The best m (first minimum) is swapped at
 *    once when the change is negative (eager);
 *  - a pass ends at the candidate of the last swap; passes repeat until one
 *    makes no swap or max_iter passes ran.  k = 1 picks the point with the
 *    least distance sum (the initial medoid unless another is strictly less).
 * Every sum over the points runs in a fixed order, the one the GPU kernels
 * use.  A candidate's sums: 64 partial sums, partial l over the points
 * o = l, l+64, l+128, ... in ascending o, then combined pairwise (l + 32 into
 * l, then l + 16 into l, ... down to l + 1).  The loss and removal losses:
 * see km_totals.  The Rust package adds in plain sequence; the orders differ
 * by rounding only.  Distances are read as D[row * n + col] with the
 * rows the package uses (mat.get(j, o) for the candidate's pass). */
#define KM_LANES 64
#define KM_NONE 0xFFFFFFFFu

typedef struct {
    uint32_t ni, si;
    double nd, sd;
} km_rec;

static double km_tree(double *p) {
    for (int s = KM_LANES / 2; s >= 1; s >>= 1)
        for (int l = 0; l < s; ++l) p[l] = p[l] + p[l + s];
    return p[0];
}

static void km_assign(const double *D, int64_t n, int k, const int64_t *med, km_rec *r) {
    for (int64_t i = 0; i < n; ++i) {
        km_rec c = {0, KM_NONE, D[i * n + med[0]], 0.0};
        for (int m = 1; m < k; ++m) {
            const double d = D[i * n + med[m]];
            if (d < c.nd || i == med[m]) {
                c.si = c.ni;
                c.sd = c.nd;
                c.ni = (uint32_t)m;
                c.nd = d;
            } else if (c.si == KM_NONE || d < c.sd) {
                c.si = (uint32_t)m;
                c.sd = d;
            }
        }
        r[i] = c;
    }
}

/* loss = sum of nearest distances; rl[m] = removal loss of medoid m.  These
 * sums run block by block: blocks of 256 points in order, each the pairwise
 * tree of its four 64-point groups' trees, (g0 + g2) + (g1 + g3), added in
 * block order to a running sum from 0. */
#define KM_BLOCK 256
static double km_totals(const km_rec *r, int64_t n, int k, double *rl, double *Q) {
    double p[KM_LANES], g[4];
    double *tot = Q; /* k + 1 running sums: removal losses, then the loss */
    for (int v = 0; v <= k; ++v) tot[v] = 0.0;
    for (int64_t b0 = 0; b0 < n; b0 += KM_BLOCK) {
        for (int v = 0; v <= k; ++v) {
            for (int w = 0; w < 4; ++w) {
                for (int t = 0; t < KM_LANES; ++t) {
                    const int64_t o = b0 + w * KM_LANES + t;
                    double x = 0.0;
                    if (o < n) x = v == k ? r[o].nd : (r[o].ni == (uint32_t)v ? r[o].sd - r[o].nd : 0.0);
                    p[t] = x;
                }
                g[w] = km_tree(p);
            }
            tot[v] += (g[0] + g[2]) + (g[1] + g[3]);
        }
    }
    for (int m = 0; m < k; ++m) rl[m] = tot[m];
    return tot[k];
}

/* loss change of the best swap that brings candidate j in; *best = the medoid out */
static double km_best_swap(const double *D, int64_t n, int k, const km_rec *r, const double *rl, int64_t j,
                           int *best, double *Q) {
    double A[KM_LANES];
    memset(Q, 0, sizeof(double) * (size_t)k * KM_LANES);
    for (int l = 0; l < KM_LANES; ++l) {
        A[l] = 0.0;
        for (int64_t o = l; o < n; o += KM_LANES) {
            const double djo = D[j * n + o];
            const km_rec c = r[o];
            if (djo < c.nd) {
                A[l] += djo - c.nd;
                Q[(size_t)c.ni * KM_LANES + l] += c.nd - c.sd;
            } else if (djo < c.sd) {
                Q[(size_t)c.ni * KM_LANES + l] += djo - c.sd;
            }
        }
    }
    const double acc = km_tree(A);
    int b = 0;
    double bl = 0.0;
    for (int m = 0; m < k; ++m) {
        const double v = rl[m] + km_tree(Q + (size_t)m * KM_LANES);
        if (m == 0 || v < bl) {
            bl = v;
            b = m;
        }
    }
    *best = b;
    return bl + acc;
}

static void km_second(const double *D, int64_t n, int k, const int64_t *med, uint32_t nn, uint32_t b, int64_t o,
                      double djo, uint32_t *si, double *sd) {
    uint32_t s = b;
    double d = djo;
    for (int i = 0; i < k; ++i) {
        if ((uint32_t)i == nn || (uint32_t)i == b) continue;
        const double e = D[o * n + med[i]];
        if (e < d) {
            s = (uint32_t)i;
            d = e;
        }
    }
    *si = s;
    *sd = d;
}

static void km_swap(const double *D, int64_t n, int k, int64_t *med, km_rec *r, uint32_t b, int64_t j) {
    med[b] = j;
    for (int64_t o = 0; o < n; ++o) {
        km_rec c = r[o];
        if (o == j) {
            if (c.ni != b) {
                c.si = c.ni;
                c.sd = c.nd;
            }
            c.ni = b;
            c.nd = 0.0;
            r[o] = c;
            continue;
        }
        const double djo = D[j * n + o];
        if (c.ni == b) {
            if (djo < c.sd) {
                c.nd = djo;
            } else {
                c.ni = c.si;
                c.nd = c.sd;
                km_second(D, n, k, med, c.ni, b, o, djo, &c.si, &c.sd);
            }
        } else if (djo < c.nd) {
            c.si = c.ni;
            c.sd = c.nd;
            c.ni = b;
            c.nd = djo;
        } else if (djo < c.sd) {
            c.si = b;
            c.sd = djo;
        } else if (c.si == b) {
            km_second(D, n, k, med, c.ni, b, o, djo, &c.si, &c.sd);
        }
        r[o] = c;
    }
}

/* sum of row j of D over o != j in the fixed order */
static double km_rowsum(const double *D, int64_t n, int64_t j) {
    double P[KM_LANES];
    for (int l = 0; l < KM_LANES; ++l) {
        P[l] = 0.0;
        for (int64_t o = l; o < n; o += KM_LANES)
            if (o != j) P[l] += D[j * n + o];
    }
    return km_tree(P);
}

/* medoids: in = initial (distinct indices), out = final.  Returns 0, or -1
 * on bad arguments / allocation failure. */
int tsmo_fasterpam(const double *D, int64_t n, int k, int64_t *med, int max_iter, int32_t *labels, double *loss_out,
                   int *iter_out, int *swaps_out) {
    if (n < 1 || k < 1 || k > n) return -1;
    if (k == 1) {
        const int64_t first = med[0];
        int64_t best = first;
        double sb = km_rowsum(D, n, first);
        for (int64_t j = 0; j < n; ++j) {
            if (j == first) continue;
            const double sj = km_rowsum(D, n, j);
            if (sj < sb) {
                best = j;
                sb = sj;
            }
        }
        med[0] = best;
        for (int64_t o = 0; o < n; ++o) labels[o] = 0;
        *loss_out = sb;
        *iter_out = 1;
        *swaps_out = best != first;
        return 0;
    }
    km_rec *r = (km_rec *)malloc(sizeof(km_rec) * (size_t)n);
    double *rl = (double *)malloc(sizeof(double) * (size_t)k);
    double *Q = (double *)malloc(sizeof(double) * ((size_t)k * KM_LANES + 1));
    if (!r || !rl || !Q) {
        free(r);
        free(rl);
        free(Q);
        return -1;
    }
    km_assign(D, n, k, med, r);
    double loss = km_totals(r, n, k, rl, Q);
    int64_t lastswap = n;
    int swaps = 0, iter = 0;
    while (iter < max_iter) {
        ++iter;
        const int before = swaps;
        for (int64_t j = 0; j < n; ++j) {
            if (j == lastswap) break;
            if (j == med[r[j].ni]) continue; /* already a medoid */
            int b;
            const double change = km_best_swap(D, n, k, r, rl, j, &b, Q);
            if (change >= 0.0) continue;
            ++swaps;
            lastswap = j;
            km_swap(D, n, k, med, r, (uint32_t)b, j);
            loss = km_totals(r, n, k, rl, Q);
        }
        if (swaps == before) break;
    }
    for (int64_t o = 0; o < n; ++o) labels[o] = (int32_t)r[o].ni;
    *loss_out = loss;
    *iter_out = iter;
    *swaps_out = swaps;
    free(r);
    free(rl);
    free(Q);
    return 0;
}

/* the loss change km_best_swap reports for candidate j in the state reached
 * from the given medoids (test hook for the swap-delta identity) */
double tsmo_fasterpam_delta(const double *D, int64_t n, int k, const int64_t *med, int64_t j, int *best) {
    km_rec *r = (km_rec *)malloc(sizeof(km_rec) * (size_t)n);
    double *rl = (double *)malloc(sizeof(double) * (size_t)k);
    double *Q = (double *)malloc(sizeof(double) * ((size_t)k * KM_LANES + 1));
    double v = NAN;
    if (r && rl && Q && k >= 2) {
        km_assign(D, n, k, med, r);
        km_totals(r, n, k, rl, Q);
        v = km_best_swap(D, n, k, r, rl, j, best, Q);
    }
    free(r);
    free(rl);
    free(Q);
    return v;
}

int tsmo_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
