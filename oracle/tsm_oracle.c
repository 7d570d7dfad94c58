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
       TSMO_STEDR = 6, TSMO_STLCSS = 7, TSMO_CDDS = 8 };

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

int tsmo_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
