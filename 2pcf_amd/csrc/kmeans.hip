// kmeans.hip -- k-means patch labels on the device: partition_data
// (src/kmeans.f90:12-263) in its Voronoi form (equal_size = .false., no
// weight equaliser: the clustering_scheme voronoi_blocks of
// src/tpcf.f90:1100-1113), which the bootstrap's lens patches come from.
//
// Rounds of: every point takes the nearest centroid (the first of equal
// minima, as the reference's strict '<' keeps it) and the moved points are
// counted; the centroids become the weighted means of their members'
// coordinates.  Stop when a round moves no point, moves fewer than tol n, or
// max_it + 1 rounds have run (the reference's loop, src/kmeans.f90:117-138).
// Start: the fixed grid of kmeans_compute_fixed_initial_centroids
// (src/kmeans.f90:85-110) with labels int(i/n k), or centroids the caller
// gives.  Spherical coordinates: great-circle distances (compared as
// chords, the same order) and plain weighted means of longitude and
// latitude, as the reference's distance_sphere and calculate_centroids do.
//
// With the weight equaliser (weight_equalizer_scaling > 0) a point's distances
// depend on cluster weights that the reference updates after every single
// move, so its result depends on the visiting order; that form is not
// offered here (TPCF_ERR_UNSUPPORTED).
//
// One thread per point, the centroids (FP64, k <= 1024) in LDS; per-block
// partial sums of w (x - x0), w (y - y0), w per cluster in LDS, added to global
// sums once per block.  The sums are 64-bit fixed point (each term rounded to
// a multiple of 2^-62 of the largest possible total), so every add is an
// integer add: the totals, the centroids and hence the labels do not depend
// on the order in which the blocks and lanes add, and a run is reproducible
// bit for bit.  A round is two launches and one 8-byte read-back.
#include <algorithm>
#include <cmath>
#include <vector>

#include "internal.hpp"

namespace tpcf {

void fill_f64(hipStream_t s, double *p, int64_t n);
void fill_u64(hipStream_t s, unsigned long long *p, int64_t n);

namespace {

constexpr int kKmBlk = 256;
constexpr int kKmMax = 1024;

// starting labels int(i / n k), src/kmeans.f90:107-108 (0-based)
__global__ __launch_bounds__(kKmBlk) void k_km_init(int64_t n, int32_t k, int32_t *__restrict__ lab)
{
    const int64_t i = blockIdx.x * (int64_t)kKmBlk + threadIdx.x;
    if (i >= n) return;
    const int32_t l = (int32_t)((double)i / (double)n * (double)k);
    lab[i] = l < k ? l : k - 1;
}

// Fixed-point frame of the centroid sums: term q of a point is rounded to an
// integer multiple of 1/scale[q] after subtracting the origin (x0, y0), with
// scale[q] = 2^62 / (n max|w| extent): no partial sum can leave the int64
// range, and each term carries an absolute error of at most 0.5 / scale[q].
struct KmFix {
    double x0, y0;
    double scale[3];   // w (x - x0), w (y - y0), w
};

__device__ __forceinline__ unsigned long long km_fix(double v, double scale)
{
    return (unsigned long long)__double2ll_rn(v * scale);   // two's complement: integer adds are exact
}

template <typename T, bool SPH>
__global__ __launch_bounds__(kKmBlk) void k_km_assign(int64_t n, int32_t k, const T *__restrict__ x,
                                                      const T *__restrict__ y, const T *__restrict__ w,
                                                      const double *__restrict__ cen, double init_best, KmFix fx,
                                                      int32_t *__restrict__ lab, unsigned long long *__restrict__ sums,
                                                      unsigned long long *__restrict__ moves)
{
    __shared__ double c0[kKmMax], c1[kKmMax], c2[kKmMax];
    __shared__ unsigned long long s[3 * kKmMax];
    __shared__ unsigned int blk_moves;
    for (int j = threadIdx.x; j < k; j += kKmBlk) {
        const double a = cen[2 * j], b = cen[2 * j + 1];
        if (SPH) {
            double sa, ca, sb, cb;
            sincos(a, &sa, &ca);
            sincos(b, &sb, &cb);
            c0[j] = cb * ca;
            c1[j] = cb * sa;
            c2[j] = sb;
        } else {
            c0[j] = a;
            c1[j] = b;
        }
        s[3 * j] = s[3 * j + 1] = s[3 * j + 2] = 0ull;
    }
    if (threadIdx.x == 0) blk_moves = 0;
    __syncthreads();
    const int64_t i = blockIdx.x * (int64_t)kKmBlk + threadIdx.x;
    if (i < n) {
        const double px = (double)x[i], py = (double)y[i], pw = (double)w[i];
        double u0 = px, u1 = py, u2 = 0.0;
        if (SPH) {
            double sa, ca, sb, cb;
            sincos(px, &sa, &ca);
            sincos(py, &sb, &cb);
            u0 = cb * ca;
            u1 = cb * sa;
            u2 = sb;
        }
        double best = init_best;
        int32_t bi = -1;
        for (int j = 0; j < k; ++j) {
            const double d0 = u0 - c0[j], d1 = u1 - c1[j];
            double d2 = fma(d0, d0, d1 * d1);
            if (SPH) {
                const double e = u2 - c2[j];
                d2 = fma(e, e, d2);
            }
            if (d2 < best) {   // first minimum kept (src/kmeans.f90:183-186); a NaN centroid never wins
                best = d2;
                bi = j;
            }
        }
        int32_t l = lab[i];
        if (bi >= 0 && bi != l) {
            l = bi;
            lab[i] = bi;
            atomicAdd(&blk_moves, 1u);
        }
        atomicAdd(&s[3 * l], km_fix((px - fx.x0) * pw, fx.scale[0]));
        atomicAdd(&s[3 * l + 1], km_fix((py - fx.y0) * pw, fx.scale[1]));
        atomicAdd(&s[3 * l + 2], km_fix(pw, fx.scale[2]));
    }
    __syncthreads();
    for (int j = threadIdx.x; j < k; j += kKmBlk)
        if (s[3 * j + 2] != 0ull || s[3 * j] != 0ull || s[3 * j + 1] != 0ull) {
            atomicAdd(&sums[3 * j], s[3 * j]);
            atomicAdd(&sums[3 * j + 1], s[3 * j + 1]);
            atomicAdd(&sums[3 * j + 2], s[3 * j + 2]);
        }
    if (threadIdx.x == 0 && blk_moves) atomicAdd(moves, (unsigned long long)blk_moves);
}

__global__ void k_km_centroids(int32_t k, const unsigned long long *__restrict__ sums, KmFix fx,
                               double *__restrict__ cen, double *__restrict__ cw)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= k) return;
    const double sw = (double)(long long)sums[3 * j + 2] / fx.scale[2];
    cw[j] = sw;
    // 0/0 = NaN for an empty cluster, as the reference
    cen[2 * j] = fx.x0 + ((double)(long long)sums[3 * j] / fx.scale[0]) / sw;
    cen[2 * j + 1] = fx.y0 + ((double)(long long)sums[3 * j + 1] / fx.scale[1]) / sw;
}

// Per block: min and max of x and y, and max |w| (out[5 blockIdx.x + q]).
template <typename T>
__global__ void k_km_bbox(int64_t n, const T *__restrict__ x, const T *__restrict__ y, const T *__restrict__ w,
                          double *__restrict__ out)
{
    __shared__ double sm[5][kKmBlk];
    double a = INFINITY, b = -INFINITY, c = INFINITY, d = -INFINITY, e = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)kKmBlk + threadIdx.x; i < n; i += (int64_t)gridDim.x * kKmBlk) {
        a = fmin(a, (double)x[i]);
        b = fmax(b, (double)x[i]);
        c = fmin(c, (double)y[i]);
        d = fmax(d, (double)y[i]);
        const double aw = fabs((double)w[i]);
        e = (aw > e || aw != aw) ? aw : e;   // a NaN weight propagates
    }
    sm[0][threadIdx.x] = a; sm[1][threadIdx.x] = b; sm[2][threadIdx.x] = c; sm[3][threadIdx.x] = d;
    sm[4][threadIdx.x] = e;
    __syncthreads();
    for (int st = kKmBlk / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            sm[0][threadIdx.x] = fmin(sm[0][threadIdx.x], sm[0][threadIdx.x + st]);
            sm[1][threadIdx.x] = fmax(sm[1][threadIdx.x], sm[1][threadIdx.x + st]);
            sm[2][threadIdx.x] = fmin(sm[2][threadIdx.x], sm[2][threadIdx.x + st]);
            sm[3][threadIdx.x] = fmax(sm[3][threadIdx.x], sm[3][threadIdx.x + st]);
            const double o = sm[4][threadIdx.x + st];
            if (o > sm[4][threadIdx.x] || o != o) sm[4][threadIdx.x] = o;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 5; ++q) out[5 * blockIdx.x + q] = sm[q][0];
}

// fixed initial centroids, src/kmeans.f90:85-110 (in FP64)
void fixed_centroids(int k, double xmin, double xmax, double ymin, double ymax, double *cen)
{
    const int n_side = (int)std::sqrt((double)k);
    if (xmax - xmin > ymax - ymin) {
        const double dy = (ymax - ymin) / n_side, dx = (xmax - xmin) * n_side / k;
        for (int j = 0; j < k; ++j) {
            const int row = (int)((dx / 2.0 + dx * j) / (xmax - xmin));
            cen[2 * j] = xmin + std::fmod(dx / 2.0 + dx * j, xmax - xmin);
            cen[2 * j + 1] = ymin + dy / 2.0 + dy * row;
        }
    } else {
        const double dx = (xmax - xmin) / n_side, dy = (ymax - ymin) * n_side / k;
        for (int j = 0; j < k; ++j) {
            const int row = (int)((dy / 2.0 + dy * j) / (ymax - ymin));
            cen[2 * j + 1] = ymin + std::fmod(dy / 2.0 + dy * j, ymax - ymin);
            cen[2 * j] = xmin + dx / 2.0 + dx * row;
        }
    }
}

template <typename T>
int kmeans(tpcf_ctx *ctx, int64_t n, const T *x, const T *y, const T *w, const tpcf_kmeans_opts *o, uint32_t flags,
           int32_t *labels, double *centroids, double *cluster_weights, int32_t *rounds, int64_t *last_moves)
{
    if (!ctx) return TPCF_ERR_ARG;
    ctx->err.clear();
    if (!o || !x || !y || !w || !labels || !centroids) return fail(ctx, TPCF_ERR_ARG, "a required argument is NULL");
    const int k = o->k;
    if (n < 1 || n >= ((int64_t)1 << 31)) return fail(ctx, TPCF_ERR_ARG, "n must be in [1, 2^31)");
    if (k < 2 || k > kKmMax) return fail(ctx, TPCF_ERR_ARG, "k must be in [2, 1024]");
    if (o->max_iterations < 0) return fail(ctx, TPCF_ERR_ARG, "max_iterations must be >= 0");
    if (o->weight_equalizer_scaling != 0.0)
        return fail(ctx, TPCF_ERR_UNSUPPORTED, "the weight-equalised k-means depends on the visiting order; "
                                               "only the Voronoi form (scaling 0) runs on the device");
    if (flags & ~(uint32_t)(TPCF_DEVICE_INPUTS | TPCF_SPHERICAL)) return fail(ctx, TPCF_ERR_ARG, "unknown flag");
    const bool dev = (flags & TPCF_DEVICE_INPUTS) != 0, sph = (flags & TPCF_SPHERICAL) != 0;
    TPCF_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;

    const unsigned nb = (unsigned)((n + kKmBlk - 1) / kKmBlk);
    const unsigned nbb = std::min<unsigned>(nb, 1024);
    const size_t in_bytes = dev ? 0 : sizeof(T) * 3 * (size_t)n;
    const size_t bytes = in_bytes + sizeof(int32_t) * n + sizeof(double) * (3 * k + 2 * k + k + 5 * nbb) + 64;
    char *wk;
    TPCF_TRY(ensure(ctx, tpcf_ctx::B_KMEANS, bytes, (void **)&wk));
    const T *dx = x, *dy = y, *dw = w;
    char *p = wk;
    if (!dev) {
        T *t = reinterpret_cast<T *>(p);
        TPCF_HIP(ctx, hipMemcpyAsync(t, x, sizeof(T) * n, hipMemcpyHostToDevice, s));
        TPCF_HIP(ctx, hipMemcpyAsync(t + n, y, sizeof(T) * n, hipMemcpyHostToDevice, s));
        TPCF_HIP(ctx, hipMemcpyAsync(t + 2 * n, w, sizeof(T) * n, hipMemcpyHostToDevice, s));
        dx = t; dy = t + n; dw = t + 2 * n;
        p += (in_bytes + 15) & ~(size_t)15;
    }
    int32_t *lab = reinterpret_cast<int32_t *>(p);
    unsigned long long *sums = reinterpret_cast<unsigned long long *>(p + ((sizeof(int32_t) * n + 15) & ~(size_t)15));
    double *cen = reinterpret_cast<double *>(sums + 3 * k), *cw = cen + 2 * k, *bb = cw + k;
    unsigned long long *moves = reinterpret_cast<unsigned long long *>(bb + 5 * nbb);

    // bounding box: the starting grid and the search bound 10 x diameter (src/kmeans.f90:62-66, :160)
    hipLaunchKernelGGL(k_km_bbox<T>, dim3(nbb), dim3(kKmBlk), 0, s, n, dx, dy, dw, bb);
    std::vector<double> hb(5 * nbb);
    TPCF_HIP(ctx, hipMemcpyAsync(hb.data(), bb, sizeof(double) * hb.size(), hipMemcpyDeviceToHost, s));
    TPCF_HIP(ctx, hipStreamSynchronize(s));
    double xmin = INFINITY, xmax = -INFINITY, ymin = INFINITY, ymax = -INFINITY, wmax = 0.0;
    for (unsigned b = 0; b < nbb; ++b) {
        xmin = std::min(xmin, hb[5 * b]); xmax = std::max(xmax, hb[5 * b + 1]);
        ymin = std::min(ymin, hb[5 * b + 2]); ymax = std::max(ymax, hb[5 * b + 3]);
        wmax = (hb[5 * b + 4] > wmax || std::isnan(hb[5 * b + 4])) ? hb[5 * b + 4] : wmax;
    }
    if (!std::isfinite(xmin) || !std::isfinite(xmax) || !std::isfinite(ymin) || !std::isfinite(ymax))
        return fail(ctx, TPCF_ERR_NONFINITE, "non-finite coordinate");
    if (!std::isfinite(wmax)) return fail(ctx, TPCF_ERR_NONFINITE, "non-finite weight");
    // the fixed-point frame: |sum| <= n max|w| extent < 2^62 for every partial sum
    KmFix fx;
    fx.x0 = xmin;
    fx.y0 = ymin;
    {
        const double wm = wmax > 0.0 ? wmax : 1.0, top = std::ldexp(1.0, 62);
        const double ex = xmax > xmin ? xmax - xmin : 1.0, ey = ymax > ymin ? ymax - ymin : 1.0;
        fx.scale[0] = top / ((double)n * wm * ex);
        fx.scale[1] = top / ((double)n * wm * ey);
        fx.scale[2] = top / ((double)n * wm);
    }
    double diam = std::sqrt((xmax - xmin) * (xmax - xmin) + (ymax - ymin) * (ymax - ymin)) * 10.0;
    // compared as squared distances (chords on the sphere: 10 diameters exceed any chord)
    const double init_best = sph ? 5.0 : diam * diam;
    std::vector<double> hc(2 * k);
    if (o->init == 1) {
        TPCF_HIP(ctx, hipMemcpyAsync(hc.data(), centroids, sizeof(double) * 2 * k,
                                     dev ? hipMemcpyDeviceToHost : hipMemcpyHostToHost, s));
        TPCF_HIP(ctx, hipStreamSynchronize(s));
    } else {
        fixed_centroids(k, xmin, xmax, ymin, ymax, hc.data());
    }
    TPCF_HIP(ctx, hipMemcpyAsync(cen, hc.data(), sizeof(double) * 2 * k, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_km_init, dim3(nb), dim3(kKmBlk), 0, s, n, k, lab);
    int it = 0;
    unsigned long long hm = 0;
    while (it <= o->max_iterations) {
        ++it;
        fill_u64(s, sums, 3 * k);
        fill_u64(s, moves, 1);
        if (sph)
            hipLaunchKernelGGL((k_km_assign<T, true>), dim3(nb), dim3(kKmBlk), 0, s, n, k, dx, dy, dw, cen, init_best,
                               fx, lab, sums, moves);
        else
            hipLaunchKernelGGL((k_km_assign<T, false>), dim3(nb), dim3(kKmBlk), 0, s, n, k, dx, dy, dw, cen, init_best,
                               fx, lab, sums, moves);
        hipLaunchKernelGGL(k_km_centroids, dim3((k + 255) / 256), dim3(256), 0, s, k, sums, fx, cen, cw);
        TPCF_HIP(ctx, hipMemcpyAsync(&hm, moves, sizeof hm, hipMemcpyDeviceToHost, s));
        TPCF_HIP(ctx, hipStreamSynchronize(s));
        if (hm == 0 || (double)hm / (double)n < o->reassignment_tol) break;
    }
    TPCF_HIP(ctx, hipGetLastError());
    const hipMemcpyKind ok = dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    TPCF_HIP(ctx, hipMemcpyAsync(labels, lab, sizeof(int32_t) * n, ok, s));
    TPCF_HIP(ctx, hipMemcpyAsync(centroids, cen, sizeof(double) * 2 * k, ok, s));
    if (cluster_weights) TPCF_HIP(ctx, hipMemcpyAsync(cluster_weights, cw, sizeof(double) * k, ok, s));
    TPCF_HIP(ctx, hipStreamSynchronize(s));
    if (rounds) *rounds = it;
    if (last_moves) *last_moves = (int64_t)hm;
    return TPCF_OK;
}

}  // namespace

}  // namespace tpcf

using namespace tpcf;

extern "C" int tpcf_kmeans_f32(tpcf_ctx *ctx, int64_t n, const float *x, const float *y, const float *w,
                               const tpcf_kmeans_opts *opts, uint32_t flags, int32_t *labels, double *centroids,
                               double *cluster_weights, int32_t *rounds, int64_t *last_moves)
{
    return kmeans<float>(ctx, n, x, y, w, opts, flags, labels, centroids, cluster_weights, rounds, last_moves);
}

extern "C" int tpcf_kmeans_f64(tpcf_ctx *ctx, int64_t n, const double *x, const double *y, const double *w,
                               const tpcf_kmeans_opts *opts, uint32_t flags, int32_t *labels, double *centroids,
                               double *cluster_weights, int32_t *rounds, int64_t *last_moves)
{
    return kmeans<double>(ctx, n, x, y, w, opts, flags, labels, centroids, cluster_weights, rounds, last_moves);
}
