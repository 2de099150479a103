/*
 * tpcf_oracle.c -- CPU FP64 oracle for 2pcf's binned pair counts.
 *
 * TEST INFRASTRUCTURE ONLY.  The product library never links, loads or calls
 * this file.  It is loaded (through oracle/oracle.py) by tests/, by
 * __graft_entry__.smoke() and by the cpu_baseline leg of bench.py, always as
 * the checker of the HIP path, never as the thing measured or shipped.
 *
 * What it restates (reference = tilmantroester/2pcf, Fortran):
 *   orc_edges            bin edges of create_bins, src/utils.f90:139-230, as an
 *                        FP64 closed form (lin / log / ihs / sqrt spacing, the
 *                        enum of src/utils.f90:7); first edge = x_min and last
 *                        edge = x_max exactly.
 *   orc_bin_of_d2        find_bin_index, src/utils.f90:101-137, on squared
 *                        separations: bin k <=> e_k^2 <= d^2 < e_{k+1}^2, with
 *                        d^2 == e_n^2 (x == x_max) going to the last bin and
 *                        anything outside [x_min, x_max] dropped.
 *   orc_correlate        the per-lens loops of tangential_shear_tpcf
 *                        (src/tpcf.f90:266-359) and scalar_kappa_tpcf
 *                        (src/tpcf.f90:456-526): masks, per-bin sums, per-lens
 *                        marks, with every pair visited as with --no-kd-tree
 *                        (src/tpcf.f90:293-295), plus per-patch-pair sums whose
 *                        sum over source patches is the per-block sum of
 *                        src/tpcf.f90:1130-1137.
 *   orc_finalize         the normalisation of src/tpcf.f90:369-376 / :536-543
 *                        (mean_theta = Theta/W with no guard, g = G/K where K>0).
 *   spherical = 1        the spherical_coords branches (src/tpcf.f90:313-320,
 *                        :337-345): x, y are longitude and latitude in radians,
 *                        the separation is the great-circle angle and the
 *                        position angle follows angles_sphere
 *                        (src/utils.f90:618-639) with FP64 trigonometry.  The
 *                        bin of a pair is decided on its squared chord
 *                        |u_A - u_B|^2 against the squared chords 2 sin(e/2) of
 *                        the edges (the same order as the angle, exact
 *                        arithmetic); the angle itself is atan2(|u_A x u_B|,
 *                        u_A . u_B), accurate at every separation (the
 *                        reference's FP32 acos of distance_sphere,
 *                        src/utils.f90:512-528, is not).  orc_sph_pair exposes
 *                        the per-pair angle and cos 2phi, sin 2phi.
 *
 * Arithmetic: coordinates arrive as double (FP32 inputs are widened exactly);
 * dx = x_A - x_B and dy = y_A - y_B are formed in FP64 and the squared
 * separation is fma(dx, dx, dy*dy).  Sums are accumulated in long double.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

enum { ORC_LIN = 0, ORC_LOG = 1, ORC_IHS = 2, ORC_SQRT = 3 };
enum { ORC_SHEAR = 0, ORC_KAPPA = 1 };

typedef struct {
    int32_t kind;                 /* ORC_SHEAR or ORC_KAPPA */
    int32_t n_bin;
    int64_t n_a, n_b;
    const double *xa, *ya, *za, *vala, *wa;          /* za unused for kappa */
    const double *xb, *yb, *zb, *s1b, *s2b, *wb, *mb; /* shear: e1,e2,m ; kappa: kappa_E,kappa_B */
    const double *edges;          /* n_bin+1 edges (not squared) */
    const int32_t *patch_a, *patch_b; /* nullable, 0-based labels */
    int32_t n_pa, n_pb;
    int32_t brute;                /* 1: every pair; 0: exact cell pruning */
    int32_t n_threads;
    int32_t spherical;            /* 1: longitude / latitude in radians (always every pair) */
} orc_args;

typedef struct {
    int64_t *n_pairs;             /* [n_bin] */
    double *theta, *gt, *gx, *w, *k;  /* raw per-bin sums [n_bin] */
    double *xi_e, *xi_b, *k_marks, *w_marks; /* nullable [n_a][n_bin], bin fastest */
    double *marks_weight;         /* nullable [n_a] */
    int64_t *pp_n;                /* nullable [n_pb][n_pa][n_bin] */
    double *pp_theta, *pp_gt, *pp_gx, *pp_w, *pp_k;
    double *gt_abs, *gx_abs;      /* nullable [n_bin]: sums of |term|, the scale of a signed sum */
} orc_out;

/* Transform of each spacing (src/utils.f90:83-90 for ihs) and its inverse. */
static double tr_fwd(int spacing, double x, double p)
{
    switch (spacing) {
    case ORC_LIN: return x;
    case ORC_LOG: return log2(x);
    case ORC_IHS: return asinh(p * x) / p;
    default: return sqrt(x);
    }
}

static double tr_inv(int spacing, double t, double p)
{
    switch (spacing) {
    case ORC_LIN: return t;
    case ORC_LOG: return exp2(t);
    case ORC_IHS: return sinh(p * t) / p;
    default: return t * t;
    }
}

/* create_bins (src/utils.f90:139-230) with x_min, x_max, n_bin given: edge i is
 * the inverse transform of T(x_min) + i*(T(x_max)-T(x_min))/n_bin.  The log
 * spacing is written with log2/exp2 and x_min factored out so that power-of-two
 * ratios give exact edges. */
int orc_edges(int spacing, float x_min, float x_max, int n_bin, float ihs_theta, double *edges)
{
    if (n_bin < 1 || !(x_max > x_min) || !(x_min >= 0.0f)) return 1;
    if (spacing < ORC_LIN || spacing > ORC_SQRT) return 1;
    if (spacing == ORC_LOG && !(x_min > 0.0f)) return 1;
    const double lo = (double)x_min, hi = (double)x_max, p = (double)ihs_theta;
    if (spacing == ORC_IHS && !(p > 0.0)) return 1;
    edges[0] = lo;
    edges[n_bin] = hi;
    if (spacing == ORC_LOG) {
        const double span = log2(hi / lo);
        for (int i = 1; i < n_bin; ++i) edges[i] = lo * exp2(((double)i * span) / (double)n_bin);
    } else {
        const double t0 = tr_fwd(spacing, lo, p), t1 = tr_fwd(spacing, hi, p);
        for (int i = 1; i < n_bin; ++i)
            edges[i] = tr_inv(spacing, t0 + ((double)i * (t1 - t0)) / (double)n_bin, p);
    }
    return 0;
}

/* Squared-edge table used by orc_bin_of_d2. */
void orc_edges_squared(const double *edges, int n_bin, double *e2)
{
    for (int i = 0; i <= n_bin; ++i) e2[i] = edges[i] * edges[i];
}

/* find_bin_index (src/utils.f90:101-137) on squared separations. */
int orc_bin_of_d2(double d2, const double *e2, int n_bin)
{
    if (d2 < e2[0] || d2 > e2[n_bin]) return -1;   /* x < x_min or x > x_max */
    if (d2 == e2[n_bin]) return n_bin - 1;          /* x == x_max -> last bin */
    int lo = 0, hi = n_bin;                          /* e2[lo] <= d2 < e2[hi] */
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (d2 >= e2[mid]) lo = mid; else hi = mid;
    }
    return lo;
}

typedef struct {
    long double theta, gt, gx, w, k, gt_abs, gx_abs;
    int64_t n;
} acc6;

static void acc_add(acc6 *dst, const acc6 *src)
{
    dst->theta += src->theta; dst->gt += src->gt; dst->gx += src->gx;
    dst->w += src->w; dst->k += src->k; dst->n += src->n;
    dst->gt_abs += src->gt_abs; dst->gx_abs += src->gx_abs;
}

/* Exact pruning structure over B: square cells no smaller than x_max, so every
 * pair with separation <= x_max lies in the 3x3 block around the lens cell. */
typedef struct {
    double x0, y0, cs;
    int64_t nx, ny;
    int64_t *start;   /* [nx*ny+1] */
    int64_t *idx;     /* [n_b] source indices sorted by cell */
} orc_grid;

static int grid_build(orc_grid *g, const orc_args *a, double x_max)
{
    double xmin = INFINITY, xmax = -INFINITY, ymin = INFINITY, ymax = -INFINITY;
    for (int64_t j = 0; j < a->n_b; ++j) {
        if (a->xb[j] < xmin) xmin = a->xb[j];
        if (a->xb[j] > xmax) xmax = a->xb[j];
        if (a->yb[j] < ymin) ymin = a->yb[j];
        if (a->yb[j] > ymax) ymax = a->yb[j];
    }
    double cs = x_max * 1.001 + 1e-300;
    const double ext = fmax(xmax - xmin, ymax - ymin);
    if (ext / cs > 4096.0) cs = ext / 4096.0;
    g->x0 = xmin; g->y0 = ymin; g->cs = cs;
    g->nx = (int64_t)floor((xmax - xmin) / cs) + 1;
    g->ny = (int64_t)floor((ymax - ymin) / cs) + 1;
    const int64_t nc = g->nx * g->ny;
    g->start = (int64_t *)calloc((size_t)nc + 1, sizeof(int64_t));
    g->idx = (int64_t *)malloc(sizeof(int64_t) * (size_t)(a->n_b > 0 ? a->n_b : 1));
    int64_t *cell = (int64_t *)malloc(sizeof(int64_t) * (size_t)(a->n_b > 0 ? a->n_b : 1));
    if (!g->start || !g->idx || !cell) { free(cell); return 1; }
    for (int64_t j = 0; j < a->n_b; ++j) {
        int64_t cx = (int64_t)floor((a->xb[j] - g->x0) / cs);
        int64_t cy = (int64_t)floor((a->yb[j] - g->y0) / cs);
        if (cx < 0) cx = 0;
        if (cx >= g->nx) cx = g->nx - 1;
        if (cy < 0) cy = 0;
        if (cy >= g->ny) cy = g->ny - 1;
        cell[j] = cy * g->nx + cx;
        g->start[cell[j] + 1]++;
    }
    for (int64_t c = 0; c < nc; ++c) g->start[c + 1] += g->start[c];
    int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (size_t)nc);
    if (!fill) { free(cell); return 1; }
    memcpy(fill, g->start, sizeof(int64_t) * (size_t)nc);
    for (int64_t j = 0; j < a->n_b; ++j) g->idx[fill[cell[j]]++] = j;
    free(fill);
    free(cell);
    return 0;
}

static void grid_free(orc_grid *g)
{
    free(g->start);
    free(g->idx);
}

/* One (lens i, source j) pair: src/tpcf.f90:302-358 (shear) / :491-525 (kappa). */
static inline void visit_pair(const orc_args *a, const double *e2, int64_t i, int64_t j,
                              acc6 *bins, acc6 *pp_row, long double *mk)
{
    const int n_bin = a->n_bin;
    const int shear = (a->kind == ORC_SHEAR);
    if (shear && (a->za[i] < 0.0 || a->zb[j] < 0.0)) return;   /* masked objects */
    const double dx = a->xa[i] - a->xb[j];
    const double dy = a->ya[i] - a->yb[j];
    const double d2 = fma(dx, dx, dy * dy);
    if (shear && a->za[i] > a->zb[j]) return;                  /* source in front of lens */
    const int b = orc_bin_of_d2(d2, e2, n_bin);
    if (b < 0) return;
    const double d = sqrt(d2);
    const double wA = a->wa[i], valA = a->vala[i], wB = a->wb[j];
    double t, x, kw;
    if (shear) {
        const double e1 = a->s1b[j], e2v = a->s2b[j], m = a->mb[j];
        const double c2 = 2.0 * (dx / d) * (dx / d) - 1.0;
        const double s2 = 2.0 * dx * dy / (d * d);
        t = -(e1 * c2 + e2v * s2) * wB * valA * wA;
        x = -(-e1 * s2 + e2v * c2) * wB * valA * wA;
        kw = wB * (1.0 + m) * wA;
    } else {
        t = a->s1b[j] * wB * valA * wA;
        x = a->s2b[j] * wB * valA * wA;
        kw = wB * wA;
    }
    const double ww = wB * wA;
    acc6 *ab = &bins[b];
    ab->theta += d * wB * wA; ab->gt += t; ab->gx += x; ab->w += ww; ab->k += kw; ab->n += 1;
    ab->gt_abs += fabs(t); ab->gx_abs += fabs(x);
    if (pp_row) {
        const int32_t pb = a->patch_b ? a->patch_b[j] : 0;
        acc6 *ap = &pp_row[(size_t)pb * n_bin + b];
        ap->theta += d * wB * wA; ap->gt += t; ap->gx += x; ap->w += ww; ap->k += kw; ap->n += 1;
    }
    if (mk) {
        long double *m4 = mk + (size_t)4 * b;
        m4[0] += t; m4[1] += x; m4[2] += kw; m4[3] += ww;
    }
}

/* Chord 2 sin(e/2) of each angular edge (angles past pi hold no pairs), squared. */
void orc_chord_edges_squared(const double *edges, int n_bin, double *e2)
{
    const double pi = 3.14159265358979323846;
    for (int i = 0; i <= n_bin; ++i) {
        const double c = 2.0 * sin(0.5 * fmin(edges[i], pi));
        e2[i] = c * c;
    }
}

static inline void unit_vec(double lon, double lat, double *u)
{
    u[0] = cos(lat) * cos(lon);
    u[1] = cos(lat) * sin(lon);
    u[2] = sin(lat);
}

/* Great-circle angle between two unit vectors, accurate at every separation. */
static inline double sph_angle(const double *u, const double *v)
{
    const double cx = u[1] * v[2] - u[2] * v[1];
    const double cy = u[2] * v[0] - u[0] * v[2];
    const double cz = u[0] * v[1] - u[1] * v[0];
    return atan2(sqrt(cx * cx + cy * cy + cz * cz), u[0] * v[0] + u[1] * v[1] + u[2] * v[2]);
}

/* angles_sphere (src/utils.f90:618-639) in FP64 and the spin-2 factors
 * cos2phi = 2 cos^2 - 1, sin2phi = 2 cos sin (src/tpcf.f90:340-342). */
static inline void sph_c2s2(double lon1, double lat1, double lon2, double lat2, double d, double *c2, double *s2)
{
    const double sin_d = sin(d);
    const double cos_lat2 = cos(lat2);
    const double ct = cos_lat2 / sin_d * sin(lon2 - lon1);
    const double st = (cos(lat1) * sin(lat2) - sin(lat1) * cos_lat2 * cos(lon2 - lon1)) / sin_d;
    *c2 = 2.0 * ct * ct - 1.0;
    *s2 = 2.0 * ct * st;
}

/* Per-pair spherical geometry for the tests: angle, cos 2phi, sin 2phi. */
void orc_sph_pair(int64_t n, const double *lon1, const double *lat1, const double *lon2, const double *lat2,
                  double *d, double *c2, double *s2)
{
    for (int64_t i = 0; i < n; ++i) {
        double u[3], v[3];
        unit_vec(lon1[i], lat1[i], u);
        unit_vec(lon2[i], lat2[i], v);
        d[i] = sph_angle(u, v);
        sph_c2s2(lon1[i], lat1[i], lon2[i], lat2[i], d[i], &c2[i], &s2[i]);
    }
}

/* One spherical (lens i, source j) pair: src/tpcf.f90:302-358 with the
 * spherical branches (:313-314, :337-342); kappa: :491-525. */
static inline void visit_pair_sph(const orc_args *a, const double *e2, int64_t i, int64_t j, const double *ua,
                                  const double *ub, acc6 *bins, acc6 *pp_row, long double *mk)
{
    const int n_bin = a->n_bin;
    const int shear = (a->kind == ORC_SHEAR);
    if (shear && (a->za[i] < 0.0 || a->zb[j] < 0.0)) return;
    if (shear && a->za[i] > a->zb[j]) return;
    const double *u = ua + 3 * i, *v = ub + 3 * j;
    const double dx = v[0] - u[0], dy = v[1] - u[1], dz = v[2] - u[2];
    const double c2 = fma(dx, dx, fma(dy, dy, dz * dz));
    const int b = orc_bin_of_d2(c2, e2, n_bin);
    if (b < 0) return;
    const double d = sph_angle(u, v);
    const double wA = a->wa[i], valA = a->vala[i], wB = a->wb[j];
    double t, x, kw;
    if (shear) {
        double cc, ss;
        sph_c2s2(a->xa[i], a->ya[i], a->xb[j], a->yb[j], d, &cc, &ss);
        const double e1 = a->s1b[j], e2v = a->s2b[j], m = a->mb[j];
        t = -(e1 * cc + e2v * ss) * wB * valA * wA;
        x = -(-e1 * ss + e2v * cc) * wB * valA * wA;
        kw = wB * (1.0 + m) * wA;
    } else {
        t = a->s1b[j] * wB * valA * wA;
        x = a->s2b[j] * wB * valA * wA;
        kw = wB * wA;
    }
    const double ww = wB * wA;
    acc6 *ab = &bins[b];
    ab->theta += d * wB * wA; ab->gt += t; ab->gx += x; ab->w += ww; ab->k += kw; ab->n += 1;
    ab->gt_abs += fabs(t); ab->gx_abs += fabs(x);
    if (pp_row) {
        const int32_t pb = a->patch_b ? a->patch_b[j] : 0;
        acc6 *ap = &pp_row[(size_t)pb * n_bin + b];
        ap->theta += d * wB * wA; ap->gt += t; ap->gx += x; ap->w += ww; ap->k += kw; ap->n += 1;
    }
    if (mk) {
        long double *m4 = mk + (size_t)4 * b;
        m4[0] += t; m4[1] += x; m4[2] += kw; m4[3] += ww;
    }
}

int orc_correlate(const orc_args *a, orc_out *o)
{
    const int n_bin = a->n_bin;
    if (n_bin < 1 || a->n_a < 0 || a->n_b < 0) return 1;
    const int want_pp = (o->pp_n != NULL);
    const int want_marks = (o->xi_e != NULL);
    const int32_t n_pa = want_pp ? (a->n_pa > 0 ? a->n_pa : 1) : 1;
    const int32_t n_pb = want_pp ? (a->n_pb > 0 ? a->n_pb : 1) : 1;
    double *e2 = (double *)malloc(sizeof(double) * (size_t)(n_bin + 1));
    if (!e2) return 2;
    const int sph = a->spherical != 0;
    const int brute = a->brute || sph;
    double *ua = NULL, *ub = NULL;
    if (sph) {
        orc_chord_edges_squared(a->edges, n_bin, e2);
        ua = (double *)malloc(sizeof(double) * 3 * (size_t)(a->n_a > 0 ? a->n_a : 1));
        ub = (double *)malloc(sizeof(double) * 3 * (size_t)(a->n_b > 0 ? a->n_b : 1));
        if (!ua || !ub) { free(ua); free(ub); free(e2); return 2; }
        for (int64_t i = 0; i < a->n_a; ++i) unit_vec(a->xa[i], a->ya[i], ua + 3 * i);
        for (int64_t j = 0; j < a->n_b; ++j) unit_vec(a->xb[j], a->yb[j], ub + 3 * j);
    } else {
        orc_edges_squared(a->edges, n_bin, e2);
    }

    orc_grid g;
    memset(&g, 0, sizeof g);
    if (!brute && grid_build(&g, a, a->edges[n_bin])) { free(e2); grid_free(&g); return 2; }

    int nt = a->n_threads > 0 ? a->n_threads : 1;
#ifndef _OPENMP
    nt = 1;
#endif
    const size_t pp_len = (size_t)n_pa * n_pb * n_bin;
    acc6 *tbins = (acc6 *)calloc((size_t)nt * n_bin, sizeof(acc6));
    acc6 *tpp = want_pp ? (acc6 *)calloc((size_t)nt * pp_len, sizeof(acc6)) : NULL;
    long double *tmk = want_marks ? (long double *)calloc((size_t)nt * 4 * n_bin, sizeof(long double)) : NULL;
    int fail = (!tbins || (want_pp && !tpp) || (want_marks && !tmk));

    if (!fail) {
#ifdef _OPENMP
#pragma omp parallel for num_threads(nt) schedule(dynamic, 16)
#endif
        for (int64_t i = 0; i < a->n_a; ++i) {
            int tid = 0;
#ifdef _OPENMP
            tid = omp_get_thread_num();
#endif
            acc6 *bins = tbins + (size_t)tid * n_bin;
            acc6 *pp_row = NULL;
            if (want_pp) {
                const int32_t pa = a->patch_a ? a->patch_a[i] : 0;
                /* rows of the [n_pb][n_pa][n_bin] table for this lens patch are
                 * strided; keep a per-thread [n_pa][n_pb][n_bin] copy instead */
                pp_row = tpp + (size_t)tid * pp_len + (size_t)pa * n_pb * n_bin;
            }
            long double *mk = NULL;
            if (want_marks) {
                mk = tmk + (size_t)tid * 4 * n_bin;
                memset(mk, 0, sizeof(long double) * 4 * (size_t)n_bin);
            }
            if (a->wa[i] != 0.0) {                     /* src/tpcf.f90:285 / :474 */
                if (sph) {
                    for (int64_t j = 0; j < a->n_b; ++j) visit_pair_sph(a, e2, i, j, ua, ub, bins, pp_row, mk);
                } else if (brute) {
                    for (int64_t j = 0; j < a->n_b; ++j) visit_pair(a, e2, i, j, bins, pp_row, mk);
                } else {
                    int64_t cx = (int64_t)floor((a->xa[i] - g.x0) / g.cs);
                    int64_t cy = (int64_t)floor((a->ya[i] - g.y0) / g.cs);
                    for (int64_t yy = cy - 1; yy <= cy + 1; ++yy) {
                        if (yy < 0 || yy >= g.ny) continue;
                        for (int64_t xx = cx - 1; xx <= cx + 1; ++xx) {
                            if (xx < 0 || xx >= g.nx) continue;
                            const int64_t c = yy * g.nx + xx;
                            for (int64_t q = g.start[c]; q < g.start[c + 1]; ++q)
                                visit_pair(a, e2, i, g.idx[q], bins, pp_row, mk);
                        }
                    }
                }
            }
            if (want_marks) {
                double mw = 0.0;
                for (int b = 0; b < n_bin; ++b) {
                    const size_t o2 = (size_t)i * n_bin + b;
                    o->xi_e[o2] = (double)mk[4 * b + 0];
                    o->xi_b[o2] = (double)mk[4 * b + 1];
                    o->k_marks[o2] = (double)mk[4 * b + 2];
                    o->w_marks[o2] = (double)mk[4 * b + 3];
                    mw += (double)mk[4 * b + 3];
                }
                if (o->marks_weight) o->marks_weight[i] = mw;
            }
        }
    }

    if (!fail) {
        for (int b = 0; b < n_bin; ++b) {
            acc6 s;
            memset(&s, 0, sizeof s);
            for (int t = 0; t < nt; ++t) acc_add(&s, &tbins[(size_t)t * n_bin + b]);
            o->n_pairs[b] = s.n;
            o->theta[b] = (double)s.theta; o->gt[b] = (double)s.gt; o->gx[b] = (double)s.gx;
            o->w[b] = (double)s.w; o->k[b] = (double)s.k;
            if (o->gt_abs) o->gt_abs[b] = (double)s.gt_abs;
            if (o->gx_abs) o->gx_abs[b] = (double)s.gx_abs;
        }
        if (want_pp) {
            for (int32_t pa = 0; pa < n_pa; ++pa)
                for (int32_t pb = 0; pb < n_pb; ++pb)
                    for (int b = 0; b < n_bin; ++b) {
                        acc6 s;
                        memset(&s, 0, sizeof s);
                        for (int t = 0; t < nt; ++t)
                            acc_add(&s, &tpp[(size_t)t * pp_len + ((size_t)pa * n_pb + pb) * n_bin + b]);
                        const size_t o3 = ((size_t)pb * n_pa + pa) * n_bin + b;
                        o->pp_n[o3] = s.n;
                        o->pp_theta[o3] = (double)s.theta; o->pp_gt[o3] = (double)s.gt;
                        o->pp_gx[o3] = (double)s.gx; o->pp_w[o3] = (double)s.w; o->pp_k[o3] = (double)s.k;
                    }
        }
    }
    free(tbins); free(tpp); free(tmk); free(e2); free(ua); free(ub);
    if (!brute) grid_free(&g);
    return fail ? 2 : 0;
}

/* Normalisation of src/tpcf.f90:369-376 (shear) / :536-543 (kappa). */
void orc_finalize(int n_bin, const double *theta, const double *gt, const double *gx,
                  const double *w, const double *k,
                  double *mean_theta, double *g_t, double *g_x)
{
    for (int b = 0; b < n_bin; ++b) {
        mean_theta[b] = theta[b] / w[b];              /* NaN for an empty bin, as the reference */
        if (k[b] > 0.0) { g_t[b] = gt[b] / k[b]; g_x[b] = gx[b] / k[b]; }
        else { g_t[b] = 0.0; g_x[b] = 0.0; }
    }
}

/* Vector form of orc_bin_of_d2 for the tests. */
void orc_bin_of_d2_array(int64_t n, const double *d2, const double *e2, int n_bin, int32_t *out)
{
    for (int64_t i = 0; i < n; ++i) out[i] = orc_bin_of_d2(d2[i], e2, n_bin);
}

/* ---- k-means patch assignment: partition_data, src/kmeans.f90:12-263 ----
 * Fixed initial centroids (kmeans_compute_fixed_initial_centroids,
 * src/kmeans.f90:85-110) and initial labels int((i-1)/n k) + 1 (here 0-based);
 * then rounds of assign_clusters (:156-200: the nearest centroid with the
 * first minimum kept, distances scaled from the second round by
 * 1 + (W_j / W_mean - 1) s when s > 0) and calculate_centroids (:202-212:
 * weighted means of the member coordinates over the tracked cluster weights)
 * until a round moves no point, moves fewer than tol n, or max_it + 1 rounds
 * have run (the loop condition of :117).  The cluster weights that scale the
 * distances move with the points: block = 1 updates them after every point,
 * as the reference's single thread does; block = B > 1 takes the points in
 * blocks of B consecutive ones, each point of a block seeing the weights from
 * before its block (B = 256: the GPU's order); block <= 0 uses the weights
 * of the previous round for the whole round.  With s = 0 all are the same.
 * equal_size = 1 (the equal_weight_blocks scheme, src/tpcf.f90:1104-1112):
 * from the second round on each point also keeps its candidate, the
 * runner-up cluster, and the ratio of the runner-up's to the winner's
 * distance (:167-180; both persist across rounds, -1 / 0 before any is set,
 * and a point at distance 0 from a centroid leaves them as they were).  After
 * the rounds, when (max W - min W) / W_mean > weight_tol, postprocess_weights
 * (:214-262): the clusters in decreasing order of weight (qsort_arg,
 * src/utils.f90:239-298, reversed), each in turn handing its members, in
 * increasing order of their ratio (qsort_arg again), to their candidate
 * cluster when that cluster comes later in the order, until its weight drops
 * to W_mean (1 + weight_tol) (checked after each hand-over).  The centroids
 * stay those of the last round.  Spherical: distances are great-circle
 * angles of (longitude, latitude) in radians and centroids the weighted means
 * of the coordinates, as the reference's distance_sphere and
 * calculate_centroids give.  FP64 throughout.  Returns the number of rounds
 * run; *pp_moves the hand-overs of the post-processing. */
static double km_dist(int sph, double x1, double y1, double x2, double y2)
{
    if (!sph) return sqrt((x1 - x2) * (x1 - x2) + (y1 - y2) * (y1 - y2));
    double u[3], v[3];
    unit_vec(x1, y1, u);
    unit_vec(x2, y2, v);
    return sph_angle(u, v);
}

/* qsort_arg (src/utils.f90:239-298): the permutation that sorts a[] in
 * increasing order, by the reference's recursive partition scheme (middle
 * pivot, scans with <= and >=, a meeting point stepped past), so that equal
 * values come out in the reference's order.  order[] is 0-based. */
static void qsort_rec(double *wk, int64_t *order, int64_t l, int64_t r)
{
    const double pivot = wk[(l + r) / 2];
    int64_t i = l - 1, j = r + 1;
    for (;;) {
        do { --j; } while (!(wk[j] <= pivot));
        do { ++i; } while (!(wk[i] >= pivot));
        if (i < j) {
            const double t = wk[i]; wk[i] = wk[j]; wk[j] = t;
            const int64_t o = order[i]; order[i] = order[j]; order[j] = o;
        } else if (i == j) {
            ++i;
            break;
        } else {
            break;
        }
    }
    if (l < j) qsort_rec(wk, order, l, j);
    if (i < r) qsort_rec(wk, order, i, r);
}

int orc_qsort_arg(int64_t n, const double *a, int64_t *order)
{
    if (n < 1) return 0;
    double *wk = (double *)malloc(sizeof(double) * (size_t)n);
    if (!wk) return -2;
    for (int64_t i = 0; i < n; ++i) { wk[i] = a[i]; order[i] = i; }
    qsort_rec(wk, order, 0, n - 1);
    free(wk);
    return 0;
}

int orc_kmeans(int64_t n, const double *x, const double *y, const double *w, int k, int max_it, double tol,
               double scaling, int block, int sph, int init_given, int equal_size, double weight_tol,
               int32_t *labels, double *cen, double *cw, int64_t *last_moves, int64_t *pp_moves)
{
    if (n < 1 || k < 2) return -1;
    double xmin = INFINITY, xmax = -INFINITY, ymin = INFINITY, ymax = -INFINITY, wsum = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        xmin = fmin(xmin, x[i]); xmax = fmax(xmax, x[i]);
        ymin = fmin(ymin, y[i]); ymax = fmax(ymax, y[i]);
        wsum += w[i];
    }
    const double diameter = sqrt((xmax - xmin) * (xmax - xmin) + (ymax - ymin) * (ymax - ymin));
    const double mean_w = wsum / k;
    if (!init_given) {
        const int n_side = (int)sqrt((double)k);
        if (xmax - xmin > ymax - ymin) {
            const double dy = (ymax - ymin) / n_side, dx = (xmax - xmin) * n_side / k;
            for (int j = 0; j < k; ++j) {
                const int row = (int)((dx / 2.0 + dx * j) / (xmax - xmin));
                cen[2 * j] = xmin + fmod(dx / 2.0 + dx * j, xmax - xmin);
                cen[2 * j + 1] = ymin + dy / 2.0 + dy * row;
            }
        } else {
            const double dx = (xmax - xmin) / n_side, dy = (ymax - ymin) * n_side / k;
            for (int j = 0; j < k; ++j) {
                const int row = (int)((dy / 2.0 + dy * j) / (ymax - ymin));
                cen[2 * j + 1] = ymin + fmod(dy / 2.0 + dy * j, ymax - ymin);
                cen[2 * j] = xmin + dx / 2.0 + dx * row;
            }
        }
    }
    for (int j = 0; j < k; ++j) cw[j] = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        int l = (int)((double)i / (double)n * k);
        labels[i] = l < k ? l : k - 1;
        cw[labels[i]] += w[i];
    }
    double *cw_prev = (double *)malloc(sizeof(double) * (size_t)k);
    double *sx = (double *)malloc(sizeof(double) * 3 * (size_t)k);
    double *mem = equal_size ? (double *)calloc((size_t)n, sizeof(double)) : NULL;
    int32_t *cand = equal_size ? (int32_t *)malloc(sizeof(int32_t) * (size_t)n) : NULL;
    if (!cw_prev || !sx || (equal_size && (!mem || !cand))) {
        free(cw_prev); free(sx); free(mem); free(cand);
        return -2;
    }
    if (equal_size)
        for (int64_t i = 0; i < n; ++i) cand[i] = -1;
    int it = 0;
    int64_t moves = 0;
    double max_dev = 0.0;
    while (it <= max_it) {
        ++it;
        moves = 0;
        const int64_t bl = block > 0 ? block : n;
        const int track = equal_size && it > 1;
        for (int64_t i = 0; i < n; ++i) {
            if (i % bl == 0) memcpy(cw_prev, cw, sizeof(double) * (size_t)k);
            const double *cwd = cw_prev;
            double best = diameter * 10.0;
            int bi = -1;
            for (int j = 0; j < k; ++j) {
                const double dw = (it > 1 && scaling > 0.0) ? 1.0 + (cwd[j] / mean_w - 1.0) * scaling : 1.0;
                const double d = dw * km_dist(sph, x[i], y[i], cen[2 * j], cen[2 * j + 1]);
                if (track && d > 0.0) {
                    if (d < best) {
                        mem[i] = best / d;
                        cand[i] = bi;
                        best = d;
                        bi = j;
                    } else if (d < mem[i] * best) {
                        mem[i] = d / best;
                        cand[i] = j;
                    }
                } else if (d < best) {
                    best = d;
                    bi = j;
                }
            }
            if (bi >= 0 && bi != labels[i]) {
                cw[labels[i]] -= w[i];
                cw[bi] += w[i];
                labels[i] = bi;
                ++moves;
            }
        }
        for (int j = 0; j < 3 * k; ++j) sx[j] = 0.0;
        for (int64_t i = 0; i < n; ++i) {
            sx[3 * labels[i]] += x[i] * w[i];
            sx[3 * labels[i] + 1] += y[i] * w[i];
            sx[3 * labels[i] + 2] += w[i];
        }
        double wmin = INFINITY, wmax = -INFINITY;
        for (int j = 0; j < k; ++j) {
            cw[j] = sx[3 * j + 2];   /* the tracked weights, re-summed exactly */
            cen[2 * j] = sx[3 * j] / cw[j];
            cen[2 * j + 1] = sx[3 * j + 1] / cw[j];
            wmin = fmin(wmin, cw[j]);
            wmax = fmax(wmax, cw[j]);
        }
        max_dev = (wmax - wmin) / mean_w;
        if (moves == 0 || (double)moves / (double)n < tol) break;
    }
    int64_t ppm = 0;
    if (equal_size && max_dev > weight_tol) {
        int64_t *ord = (int64_t *)malloc(sizeof(int64_t) * (size_t)k);
        int64_t *cord = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
        char *done = (char *)calloc((size_t)k, 1);
        int64_t *memb = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
        if (!ord || !cord || !done || !memb || orc_qsort_arg(k, cw, ord) || orc_qsort_arg(n, mem, cord)) {
            free(ord); free(cord); free(done); free(memb); free(cw_prev); free(sx); free(mem); free(cand);
            return -2;
        }
        for (int a = 0, b = k - 1; a < b; ++a, --b) {   /* reverse = .true. */
            const int64_t t = ord[a]; ord[a] = ord[b]; ord[b] = t;
        }
        const double limit = mean_w * (1.0 + weight_tol);
        for (int jj = 0; jj < k; ++jj) {
            const int c = (int)ord[jj];
            done[c] = 1;   /* cluster_size_order(1:j) holds the clusters up to this one */
            /* the members in candidate order, packed before any is handed over */
            int64_t nm = 0;
            for (int64_t q = 0; q < n; ++q)
                if (labels[cord[q]] == c) memb[nm++] = cord[q];
            for (int64_t q = 0; q < nm; ++q) {
                const int64_t p = memb[q];
                const int dst = cand[p];
                if (dst < 0 || done[dst]) continue;
                labels[p] = dst;
                cw[c] -= w[p];
                cw[dst] += w[p];
                ++ppm;
                if (cw[c] <= limit) break;
            }
        }
        free(ord); free(cord); free(done); free(memb);
    }
    free(cw_prev);
    free(sx);
    free(mem);
    free(cand);
    if (last_moves) *last_moves = moves;
    if (pp_moves) *pp_moves = ppm;
    return it;
}
