/*
 * ref_port.c -- host-core timing port of the reference's correlation call.
 *
 * TEST / BENCHMARK INFRASTRUCTURE ONLY.  This is the CPU baseline that bench.py
 * times beside the GPU path (cpu_baseline.kind = "port"); the product never
 * links it.  It restates the reference algorithm in C with the reference's
 * precision (Fortran default real = FP32) and data flow:
 *   rp_create_bins   create_bins with x_min/x_max/n_bin, src/utils.f90:139-230 (FP32)
 *   rp_find_bin      find_bin_index, src/utils.f90:101-137 (FP32 log10 etc.)
 *   kd-tree          grow_tree (midpoint split of the widest bbox side, leaf size
 *                    10), src/kdtree.f90:73-205, and search_within_r with whole-
 *                    node inclusion and box pruning, src/kdtree.f90:207-375
 *   rp_correlate     tangential_shear_tpcf src/tpcf.f90:190-378 and
 *                    scalar_kappa_tpcf src/tpcf.f90:380-545: OpenMP over
 *                    lenses, FP32 per-thread reduction arrays, per-lens marks.
 * Differences kept on purpose: the per-thread neighbour buffer grows instead of
 * overflowing (src/kdtree.f90:314-317 writes past it), and there is no
 * per-lens progress critical section (src/tpcf.f90:267-276, verbose only).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

enum { RP_LIN = 0, RP_LOG = 1, RP_IHS = 2, RP_SQRT = 3 };

typedef struct {
    float x_min, x_max, bin_size, ihs_theta;
    int32_t n_bin, spacing;
} rp_bins;

static float rp_ihs(float x, float p) { return logf(p * x + sqrtf((p * x) * (p * x) + 1.0f)) / p; }

int rp_create_bins(float x_min, float x_max, int n_bin, int spacing, float ihs_theta, rp_bins *b)
{
    if (n_bin < 1 || spacing < RP_LIN || spacing > RP_SQRT) return 1;
    b->x_min = x_min; b->x_max = x_max; b->n_bin = n_bin; b->spacing = spacing;
    b->ihs_theta = (spacing == RP_IHS) ? ihs_theta : 1.0f;
    switch (spacing) {
    case RP_LIN: b->bin_size = (x_max - x_min) / (float)n_bin; break;
    case RP_LOG: b->bin_size = (log10f(x_max) - log10f(x_min)) / (float)n_bin; break;
    case RP_IHS: b->bin_size = (rp_ihs(x_max, b->ihs_theta) - rp_ihs(x_min, b->ihs_theta)) / (float)n_bin; break;
    default: b->bin_size = (sqrtf(x_max) - sqrtf(x_min)) / (float)n_bin; break;
    }
    return 0;
}

/* 1-based bin as the reference returns it, -1 when outside. */
int rp_find_bin(float x, const rp_bins *b)
{
    if (x < b->x_min || x > b->x_max) return -1;
    if (x == b->x_max) return b->n_bin;
    float t, t0;
    switch (b->spacing) {
    case RP_LIN: t = x; t0 = b->x_min; break;
    case RP_LOG: t = log10f(x); t0 = log10f(b->x_min); break;
    case RP_IHS: t = rp_ihs(x, b->ihs_theta); t0 = rp_ihs(b->x_min, b->ihs_theta); break;
    default: t = sqrtf(x); t0 = sqrtf(b->x_min); break;
    }
    int i = (int)((t - t0) / b->bin_size) + 1;
    if (i > b->n_bin) i = -1;
    return i;
}

/* ---- kd-tree over the sources ---- */
typedef struct {
    int32_t l, u;          /* inclusive index range into perm, l > u for an empty node */
    int32_t left, right;   /* child nodes, -1 for a leaf */
    float lo[2], hi[2];    /* bounding box */
} rp_node;

typedef struct {
    const float *x, *y;
    int32_t *perm;
    rp_node *nodes;
    int32_t n_nodes, cap, leafsize;
} rp_tree;

static float coord(const rp_tree *t, int32_t k, int dim) { return dim == 0 ? t->x[k] : t->y[k]; }

static int32_t new_node(rp_tree *t)
{
    if (t->n_nodes == t->cap) {
        int32_t nc = t->cap ? t->cap * 2 : 1024;
        rp_node *nn = (rp_node *)realloc(t->nodes, sizeof(rp_node) * (size_t)nc);
        if (!nn) return -1;
        t->nodes = nn; t->cap = nc;
    }
    return t->n_nodes++;
}

static int32_t build(rp_tree *t, int32_t l, int32_t u, const float lo_in[2], const float hi_in[2])
{
    int32_t id = new_node(t);
    if (id < 0) return -1;
    rp_node nd;
    nd.l = l; nd.u = u; nd.left = nd.right = -1;
    nd.lo[0] = lo_in[0]; nd.lo[1] = lo_in[1]; nd.hi[0] = hi_in[0]; nd.hi[1] = hi_in[1];
    if (u - l + 1 <= t->leafsize) {
        if (u >= l) {                          /* leaf: tight box of its points */
            for (int d = 0; d < 2; ++d) { nd.lo[d] = INFINITY; nd.hi[d] = -INFINITY; }
            for (int32_t k = l; k <= u; ++k)
                for (int d = 0; d < 2; ++d) {
                    float v = coord(t, t->perm[k], d);
                    if (v < nd.lo[d]) nd.lo[d] = v;
                    if (v > nd.hi[d]) nd.hi[d] = v;
                }
        } else {
            nd.l = 0; nd.u = -1;               /* empty node */
        }
        t->nodes[id] = nd;
        return id;
    }
    const int dim = (hi_in[1] - lo_in[1] > hi_in[0] - lo_in[0]) ? 1 : 0;
    const float mid = (lo_in[dim] + hi_in[dim]) / 2.0f;
    int32_t a = l, b = u;                      /* partition: <= mid to the left */
    while (a < b) {
        if (coord(t, t->perm[a], dim) <= mid) {
            ++a;
        } else {
            int32_t tmp = t->perm[a]; t->perm[a] = t->perm[b]; t->perm[b] = tmp;
            --b;
        }
    }
    if (coord(t, t->perm[a], dim) > mid) --a;
    float llo[2] = {lo_in[0], lo_in[1]}, lhi[2] = {hi_in[0], hi_in[1]};
    float rlo[2] = {lo_in[0], lo_in[1]}, rhi[2] = {hi_in[0], hi_in[1]};
    lhi[dim] = mid; rlo[dim] = mid;
    if (l > a) { llo[0] = rlo[0]; llo[1] = rlo[1]; lhi[0] = rhi[0]; lhi[1] = rhi[1]; }
    else if (a + 1 > u) { rlo[0] = llo[0]; rlo[1] = llo[1]; rhi[0] = lhi[0]; rhi[1] = lhi[1]; }
    int32_t left = build(t, l, a, llo, lhi);
    int32_t right = build(t, a + 1, u, rlo, rhi);
    if (left < 0 || right < 0) return -1;
    nd.left = left; nd.right = right;
    for (int d = 0; d < 2; ++d) {
        nd.lo[d] = fminf(t->nodes[left].lo[d], t->nodes[right].lo[d]);
        nd.hi[d] = fmaxf(t->nodes[left].hi[d], t->nodes[right].hi[d]);
    }
    t->nodes[id] = nd;
    return id;
}

static int tree_grow(rp_tree *t, const float *x, const float *y, int32_t n, int32_t leafsize)
{
    memset(t, 0, sizeof *t);
    t->x = x; t->y = y; t->leafsize = leafsize;
    t->perm = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    if (!t->perm) return 1;
    for (int32_t k = 0; k < n; ++k) t->perm[k] = k;
    float lo[2] = {INFINITY, INFINITY}, hi[2] = {-INFINITY, -INFINITY};
    for (int32_t k = 0; k < n; ++k) {
        if (x[k] < lo[0]) lo[0] = x[k];
        if (x[k] > hi[0]) hi[0] = x[k];
        if (y[k] < lo[1]) lo[1] = y[k];
        if (y[k] > hi[1]) hi[1] = y[k];
    }
    return build(t, 0, n - 1, lo, hi) < 0;
}

static void tree_burn(rp_tree *t) { free(t->perm); free(t->nodes); }

typedef struct { int32_t *v; int64_t n, cap; } rp_list;

static void push(rp_list *r, int32_t k)
{
    if (r->n == r->cap) {
        r->cap = r->cap ? r->cap * 2 : 4096;
        r->v = (int32_t *)realloc(r->v, sizeof(int32_t) * (size_t)r->cap);
    }
    r->v[r->n++] = k;
}

static void search(const rp_tree *t, int32_t id, float px, float py, float r, rp_list *out)
{
    const rp_node *nd = &t->nodes[id];
    if (nd->u < nd->l) return;
    float dmin = 0.0f, dmax = 0.0f;
    const float p[2] = {px, py};
    for (int d = 0; d < 2; ++d) {
        float dl = nd->lo[d] - p[d], dr = nd->hi[d] - p[d];
        float dl2 = dl * dl, dr2 = dr * dr;
        dmax += dl2 > dr2 ? dl2 : dr2;
        if (!(dl < 0.0f && dr > 0.0f)) dmin += dl2 > dr2 ? dr2 : dl2;
    }
    if (sqrtf(dmin) > r) return;
    if (sqrtf(dmax) <= r) {
        for (int32_t k = nd->l; k <= nd->u; ++k) push(out, t->perm[k]);
        return;
    }
    if (nd->left < 0) {
        for (int32_t k = nd->l; k <= nd->u; ++k) {
            int32_t j = t->perm[k];
            float dx = t->x[j] - px, dy = t->y[j] - py;
            if (dx * dx + dy * dy <= r * r) push(out, j);
        }
        return;
    }
    search(t, nd->left, px, py, r, out);
    search(t, nd->right, px, py, r, out);
}

typedef struct {
    int32_t kind;              /* 0 shear, 1 kappa */
    int32_t n_a, n_b;
    const float *xa, *ya, *za, *vala, *wa;
    const float *xb, *yb, *zb, *s1b, *s2b, *wb, *mb;
    rp_bins bins;
    int32_t n_threads;
    int32_t use_tree;
} rp_args;

typedef struct {
    float *mean_theta, *g_t, *g_x, *weights;  /* [n_bin] */
    int64_t *n_pairs;
    float *xi_e, *xi_b, *k_marks, *w_marks, *marks_weight; /* [n_a][n_bin] / [n_a], nullable */
    double tree_seconds;
} rp_out;

int rp_correlate(const rp_args *a, rp_out *o)
{
    const int nb = a->bins.n_bin;
    const int shear = (a->kind == 0);
    rp_tree tree;
    memset(&tree, 0, sizeof tree);
    double t0 = 0.0;
#ifdef _OPENMP
    t0 = omp_get_wtime();
#endif
    if (a->use_tree && tree_grow(&tree, a->xb, a->yb, a->n_b, 10)) return 2;
#ifdef _OPENMP
    o->tree_seconds = omp_get_wtime() - t0;
#else
    o->tree_seconds = 0.0;
#endif
    int nt = a->n_threads > 0 ? a->n_threads : 1;
    float *sum = (float *)calloc((size_t)nt * 6 * nb, sizeof(float));
    int64_t *cnt = (int64_t *)calloc((size_t)nt * nb, sizeof(int64_t));
    if (!sum || !cnt) { free(sum); free(cnt); tree_burn(&tree); return 2; }
    const int marks = (o->xi_e != NULL);
    if (marks) {
        memset(o->xi_e, 0, sizeof(float) * (size_t)a->n_a * nb);
        memset(o->xi_b, 0, sizeof(float) * (size_t)a->n_a * nb);
        memset(o->k_marks, 0, sizeof(float) * (size_t)a->n_a * nb);
        memset(o->w_marks, 0, sizeof(float) * (size_t)a->n_a * nb);
        memset(o->marks_weight, 0, sizeof(float) * (size_t)a->n_a);
    }
#ifdef _OPENMP
#pragma omp parallel num_threads(nt)
#endif
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        float *s = sum + (size_t)tid * 6 * nb;   /* theta, gt, gx, w, K per bin */
        int64_t *c = cnt + (size_t)tid * nb;
        rp_list found = {0, 0, 0};
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 64)
#endif
        for (int32_t i = 0; i < a->n_a; ++i) {
            const float xf = a->xa[i], yf = a->ya[i], vf = a->vala[i], wf = a->wa[i];
            const float zf = shear ? a->za[i] : 0.0f;
            if (wf == 0.0f) continue;
            int64_t nfound;
            found.n = 0;
            if (a->use_tree) {
                search(&tree, 0, xf, yf, a->bins.x_max, &found);
                nfound = found.n;
            } else {
                nfound = a->n_b;
            }
            for (int64_t q = 0; q < nfound; ++q) {
                const int32_t j = a->use_tree ? found.v[q] : (int32_t)q;
                if (shear && (zf < 0.0f || a->zb[j] < 0.0f)) continue;
                const float dx = xf - a->xb[j], dy = yf - a->yb[j];
                const float d = sqrtf(dx * dx + dy * dy);
                if (shear && zf > a->zb[j]) continue;
                const int b1 = rp_find_bin(d, &a->bins);
                if (b1 < 1) continue;
                const int b = b1 - 1;
                const float w = a->wb[j];
                float t, x, k;
                if (shear) {
                    const float e1 = a->s1b[j], e2 = a->s2b[j], m = a->mb[j];
                    const float c2 = 2.0f * (dx / d) * (dx / d) - 1.0f;
                    const float s2 = 2.0f * dx * dy / (d * d);
                    t = (e1 * c2 + e2 * s2) * w * vf * wf;
                    x = (-e1 * s2 + e2 * c2) * w * vf * wf;
                    s[6 * b + 1] -= t; s[6 * b + 2] -= x;
                    k = w * (1.0f + m) * wf;
                    if (marks) { o->xi_e[(size_t)i * nb + b] -= t; o->xi_b[(size_t)i * nb + b] -= x; }
                } else {
                    t = a->s1b[j] * w * vf * wf;
                    x = a->s2b[j] * w * vf * wf;
                    s[6 * b + 1] += t; s[6 * b + 2] += x;
                    k = w * wf;
                    if (marks) { o->xi_e[(size_t)i * nb + b] += t; o->xi_b[(size_t)i * nb + b] += x; }
                }
                s[6 * b + 0] += d * w * wf;
                s[6 * b + 3] += w * wf;
                s[6 * b + 4] += k;
                c[b] += 1;
                if (marks) {
                    o->w_marks[(size_t)i * nb + b] += w * wf;
                    o->k_marks[(size_t)i * nb + b] += k;
                    o->marks_weight[i] += w * wf;
                }
            }
        }
        free(found.v);
    }
    for (int b = 0; b < nb; ++b) {
        float th = 0, gt = 0, gx = 0, w = 0, k = 0;
        int64_t n = 0;
        for (int t = 0; t < nt; ++t) {
            const float *s = sum + (size_t)t * 6 * nb;
            th += s[6 * b]; gt += s[6 * b + 1]; gx += s[6 * b + 2]; w += s[6 * b + 3]; k += s[6 * b + 4];
            n += cnt[(size_t)t * nb + b];
        }
        o->mean_theta[b] = th / w;
        o->g_t[b] = k > 0.0f ? gt / k : 0.0f;
        o->g_x[b] = k > 0.0f ? gx / k : 0.0f;
        o->weights[b] = w;
        o->n_pairs[b] = n;
    }
    free(sum); free(cnt);
    if (a->use_tree) tree_burn(&tree);
    return 0;
}

/* Vector form of rp_find_bin for the tests (1-based bins, -1 outside). */
void rp_find_bin_array(int64_t n, const float *x, const rp_bins *b, int32_t *out)
{
    for (int64_t i = 0; i < n; ++i) out[i] = rp_find_bin(x[i], b);
}
