// kmeans.hip -- k-means patch labels on the device: partition_data
// (src/kmeans.f90:12-263), the clustering of the voronoi_blocks and
// equal_weight_blocks schemes (src/tpcf.f90:1100-1113) that the bootstrap's
// lens patches come from.
//
// Rounds of: every point takes the nearest centroid (the first of equal
// minima, as the reference's strict '<' keeps it) and the moved points are
// counted; the centroids become the weighted means of their members'
// coordinates.  Stop when a round moves no point, moves fewer than tol n, or
// max_it + 1 rounds have run (the reference's loop, src/kmeans.f90:117-138).
// Start: the fixed grid of kmeans_compute_fixed_initial_centroids
// (src/kmeans.f90:85-110) with labels int(i/n k), or centroids the caller
// gives.  Spherical coordinates: great-circle distances (compared as
// chords, the same order, in the Voronoi form) and plain weighted means of
// longitude and latitude, as the reference's distance_sphere and
// calculate_centroids do.
//
// Weight equaliser (weight_equalizer_scaling s > 0): from the second round a
// distance to cluster j is scaled by 1 + (W_j / W_mean - 1) s.  The reference
// moves W_j point by point while a round runs, so its result depends on the
// visiting order.  Here one workgroup takes the points in steps of 256
// consecutive ones: every point of a step sees the weights from before the
// step, and the step's moves update them before the next step, so the result
// does not depend on any thread order (k_km_assign_gs).  Equal size
// (equal_size): from the second round each point also keeps its runner-up
// cluster and the ratio of the two distances (src/kmeans.f90:167-180), and
// after the rounds, when (max W - min W) / W_mean > weight_tol,
// postprocess_weights (src/kmeans.f90:214-262) runs on the device: clusters
// in decreasing order of weight (the reference's qsort_arg, restated on the
// host for the k weights), each handing its members in increasing order of
// their ratio (a stable radix sort: equal ratios by point index, where the
// reference's quicksort has its own order) to their runner-up if that cluster
// comes later, until its weight drops to W_mean (1 + weight_tol).  A member
// handed over keeps its runner-up as its cluster, so when that cluster's turn
// comes it is skipped there; each cluster's turn therefore only reads its
// members at the start of the post-processing (one segment of the sorted
// points) and the weight it has received, which one workgroup carries from
// turn to turn.
//
// One thread per point, the centroids (FP64, k <= 1024) in LDS; per-block
// partial sums of w (x - x0), w (y - y0), w per cluster in LDS, added to global
// sums once per block.  The sums are 64-bit fixed point (each term rounded to
// a multiple of 2^-62 of the largest possible total for the coordinates, and
// to a power of two at most 2^-52 of it for the weights, which keeps the sums
// of single-precision weights exact), so every add is an integer add: the
// totals, the centroids and hence the labels do not depend on the order in
// which the blocks and lanes add, and a run is reproducible bit for bit.  A
// round is an assignment launch (one workgroup with the equaliser, one thread
// per point otherwise), a centroid launch and one 8-byte read-back.
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <climits>
#include <cmath>
#include <vector>

#include "internal.hpp"

namespace tpcf {

void fill_f64(hipStream_t s, double *p, int64_t n);
void fill_u64(hipStream_t s, unsigned long long *p, int64_t n);

namespace {

constexpr int kKmBlk = 256;
constexpr int kKmMax = 1024;

__global__ __launch_bounds__(kKmBlk) void k_km_fill_i32(int64_t n, int32_t *__restrict__ p, int32_t v)
{
    const int64_t i = blockIdx.x * (int64_t)kKmBlk + threadIdx.x;
    if (i < n) p[i] = v;
}

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

// EQ: actual (scaled) distances, for the weight equaliser and the equal-size
// runner-up tracking; dwt[j] the distance scale of cluster j (NULL: 1), mem /
// cand the runner-up ratio and cluster of each point (NULL: not tracked).
template <typename T, bool SPH, bool EQ>
__global__ __launch_bounds__(kKmBlk) void k_km_assign(int64_t n, int32_t k, const T *__restrict__ x,
                                                      const T *__restrict__ y, const T *__restrict__ w,
                                                      const double *__restrict__ cen, double init_best, KmFix fx,
                                                      int32_t *__restrict__ lab, unsigned long long *__restrict__ sums,
                                                      unsigned long long *__restrict__ moves,
                                                      const double *__restrict__ dwt, double *__restrict__ mem,
                                                      int32_t *__restrict__ cand)
{
    __shared__ double c0[kKmMax], c1[kKmMax], c2[kKmMax];
    __shared__ double dws[EQ ? kKmMax : 1];
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
        if (EQ) dws[j] = dwt ? dwt[j] : 1.0;
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
        double m_i = 0.0;
        int32_t c_i = -1;
        const bool track = EQ && mem != nullptr;
        if (track) {
            m_i = mem[i];
            c_i = cand[i];
        }
        for (int j = 0; j < k; ++j) {
            if (!EQ) {
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
            } else {
                double dist;
                if (SPH) {   // the angle from the cross and dot products: accurate at every separation
                    const double cx = u1 * c2[j] - u2 * c1[j], cy = u2 * c0[j] - u0 * c2[j],
                                 cz = u0 * c1[j] - u1 * c0[j];
                    dist = atan2(sqrt(cx * cx + cy * cy + cz * cz), u0 * c0[j] + u1 * c1[j] + u2 * c2[j]);
                } else {
                    const double d0 = px - c0[j], d1 = py - c1[j];
                    dist = sqrt(d0 * d0 + d1 * d1);
                }
                const double d = dws[j] * dist;
                if (track && d > 0.0) {   // runner-up kept (src/kmeans.f90:167-180)
                    if (d < best) {
                        m_i = best / d;
                        c_i = bi;
                        best = d;
                        bi = j;
                    } else if (d < m_i * best) {
                        m_i = d / best;
                        c_i = j;
                    }
                } else if (d < best) {
                    best = d;
                    bi = j;
                }
            }
        }
        if (track) {
            mem[i] = m_i;
            cand[i] = c_i;
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

// A round of the weight equaliser (second round on): the cluster weights
// that scale the distances, 1 + (W_j / W_mean - 1) s, move with the points as
// the reference's do (src/kmeans.f90:162-200), but a block of kKmGs
// consecutive points at a time: one workgroup takes the points in steps of
// kKmGs, every point of a step sees the weights from before its step, and the
// step's moves update the weights (exact fixed-point integers in LDS) before
// the next.  The result does not depend on any thread order.  Taking the
// points one at a time, as the reference's single thread does, is a step of
// one point; the previous round's weights for the whole round made the
// clusters swing (all points leave an overweight cluster at once).  cw: the
// weights at the start of the round (FP64 of the fixed-point sums, exact).
// Also tracks the runner-ups (mem / cand, NULL: not tracked) and adds the
// round's centroid sums.
constexpr int kKmGs = 256;

template <typename T, bool SPH>
__global__ __launch_bounds__(kKmGs) void k_km_assign_gs(int64_t n, int32_t k, const T *__restrict__ x,
                                                        const T *__restrict__ y, const T *__restrict__ w,
                                                        const double *__restrict__ cen, double init_best, KmFix fx,
                                                        const double *__restrict__ cw, double scaling,
                                                        int32_t *__restrict__ lab,
                                                        unsigned long long *__restrict__ sums,
                                                        unsigned long long *__restrict__ moves,
                                                        double *__restrict__ mem, int32_t *__restrict__ cand)
{
    __shared__ double c0[kKmMax], c1[kKmMax], c2[kKmMax], dws[kKmMax];
    __shared__ unsigned long long s[3 * kKmMax];
    __shared__ unsigned long long wl[kKmMax];
    __shared__ unsigned long long tot;
    const int t = threadIdx.x;
    if (t == 0) tot = 0ull;
    __syncthreads();
    long long part = 0;
    for (int j = t; j < k; j += kKmGs) {
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
        const long long wj = (long long)(cw[j] * fx.scale[2]);   // exact: cw is a fixed-point sum
        wl[j] = (unsigned long long)wj;
        part += wj;
        s[3 * j] = s[3 * j + 1] = s[3 * j + 2] = 0ull;
    }
    if (part) atomicAdd(&tot, (unsigned long long)part);
    __syncthreads();
    const double mean = (double)(long long)tot / fx.scale[2] / (double)k;
    for (int j = t; j < k; j += kKmGs) dws[j] = 1.0 + ((double)(long long)wl[j] / fx.scale[2] / mean - 1.0) * scaling;
    __syncthreads();
    unsigned long long my_moves = 0;
    const bool track = mem != nullptr;
    for (int64_t base = 0; base < n; base += kKmGs) {
        const int64_t i = base + t;
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
            double m_i = 0.0;
            int32_t c_i = -1;
            if (track) {
                m_i = mem[i];
                c_i = cand[i];
            }
            for (int j = 0; j < k; ++j) {
                double dist;
                if (SPH) {
                    const double cx = u1 * c2[j] - u2 * c1[j], cy = u2 * c0[j] - u0 * c2[j],
                                 cz = u0 * c1[j] - u1 * c0[j];
                    dist = atan2(sqrt(cx * cx + cy * cy + cz * cz), u0 * c0[j] + u1 * c1[j] + u2 * c2[j]);
                } else {
                    const double d0 = px - c0[j], d1 = py - c1[j];
                    dist = sqrt(d0 * d0 + d1 * d1);
                }
                const double d = dws[j] * dist;
                if (track && d > 0.0) {
                    if (d < best) {
                        m_i = best / d;
                        c_i = bi;
                        best = d;
                        bi = j;
                    } else if (d < m_i * best) {
                        m_i = d / best;
                        c_i = j;
                    }
                } else if (d < best) {
                    best = d;
                    bi = j;
                }
            }
            if (track) {
                mem[i] = m_i;
                cand[i] = c_i;
            }
            int32_t l = lab[i];
            const unsigned long long wf = km_fix(pw, fx.scale[2]);
            if (bi >= 0 && bi != l) {
                atomicAdd(&wl[l], (unsigned long long)(-(long long)wf));
                atomicAdd(&wl[bi], wf);
                l = bi;
                lab[i] = bi;
                ++my_moves;
            }
            atomicAdd(&s[3 * l], km_fix((px - fx.x0) * pw, fx.scale[0]));
            atomicAdd(&s[3 * l + 1], km_fix((py - fx.y0) * pw, fx.scale[1]));
            atomicAdd(&s[3 * l + 2], wf);
        }
        __syncthreads();
        for (int j = t; j < k; j += kKmGs)
            dws[j] = 1.0 + ((double)(long long)wl[j] / fx.scale[2] / mean - 1.0) * scaling;
        __syncthreads();
    }
    for (int j = t; j < k; j += kKmGs) {
        sums[3 * j] = s[3 * j];
        sums[3 * j + 1] = s[3 * j + 1];
        sums[3 * j + 2] = s[3 * j + 2];
    }
    if (my_moves) atomicAdd(moves, my_moves);
}

// Keys of the post-processing sort: every point's runner-up ratio, and the
// turn (rank) of its cluster when it may be handed over (a runner-up whose
// turn comes later), else k; counts[r]: such points of the cluster of turn r.
__global__ __launch_bounds__(kKmBlk) void k_km_pp_keys(int64_t n, int32_t k, const int32_t *__restrict__ lab,
                                                       const int32_t *__restrict__ cand,
                                                       const int32_t *__restrict__ rank, uint32_t *__restrict__ key,
                                                       uint32_t *__restrict__ idx, uint32_t *__restrict__ counts)
{
    const int64_t i = blockIdx.x * (int64_t)kKmBlk + threadIdx.x;
    if (i >= n) return;
    const int32_t c = cand[i], l = lab[i];
    const bool elig = c >= 0 && rank[c] > rank[l];
    key[i] = elig ? (uint32_t)rank[l] : (uint32_t)k;
    idx[i] = (uint32_t)i;
    if (elig) atomicAdd(&counts[rank[l]], 1u);
}

__global__ __launch_bounds__(kKmBlk) void k_km_gather_u32(int64_t n, const uint32_t *__restrict__ idx,
                                                          const uint32_t *__restrict__ in, uint32_t *__restrict__ out)
{
    const int64_t i = blockIdx.x * (int64_t)kKmBlk + threadIdx.x;
    if (i < n) out[i] = in[idx[i]];
}

// Inclusive prefix sum over the 1024 threads of the block; *total gets the sum.
__device__ long long block_scan_incl(long long v, long long *wsum, long long *total)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int d = 1; d < 64; d <<= 1) {
        const long long o = __shfl_up(v, d, 64);
        if (lane >= d) v += o;
    }
    if (lane == 63) wsum[wv] = v;
    __syncthreads();
    if (threadIdx.x < 16) {
        long long t = wsum[threadIdx.x];
        for (int d = 1; d < 16; d <<= 1) {
            const long long o = __shfl_up(t, d, 16);
            if ((int)threadIdx.x >= d) t += o;
        }
        wsum[threadIdx.x] = t;   // inclusive over the waves
    }
    __syncthreads();
    const long long r = v + (wv > 0 ? wsum[wv - 1] : 0ll);
    *total = wsum[15];
    __syncthreads();
    return r;
}

// postprocess_weights (src/kmeans.f90:214-262) in one workgroup, turn by
// turn: cluster c = order[r] starts from its last-round weight plus what
// earlier turns handed to it, and hands its candidates (pidx[off[r] ..
// off[r+1]), increasing ratio) to their runner-ups, the first m of them with m
// the smallest count >= 1 after which its weight is <= limit (all of them if
// none).  Weights in the fixed point of the round sums (exact integers).
template <typename T>
__global__ __launch_bounds__(1024) void k_km_pp_seq(int32_t k, const int32_t *__restrict__ order,
                                                    const uint32_t *__restrict__ counts,
                                                    const uint32_t *__restrict__ pidx, int32_t *__restrict__ lab,
                                                    const int32_t *__restrict__ cand, const T *__restrict__ w,
                                                    double wscale, const unsigned long long *__restrict__ sums,
                                                    long long limit, double *__restrict__ cw,
                                                    unsigned long long *__restrict__ ppm)
{
    __shared__ unsigned long long recv[kKmMax];
    __shared__ uint32_t off[kKmMax + 1];
    __shared__ long long wsum[16];
    __shared__ int first;
    const int t = threadIdx.x;
    recv[t] = 0ull;
    {
        long long tot;
        const long long c = t < k ? (long long)counts[t] : 0ll;
        const long long incl = block_scan_incl(c, wsum, &tot);
        off[t + 1] = (uint32_t)incl;
        if (t == 0) off[0] = 0;
    }
    __syncthreads();
    unsigned long long moved_n = 0;
    for (int r = 0; r < k; ++r) {
        const int c = order[r];
        const uint32_t s0 = off[r], s1 = off[r + 1];
        const long long W = (long long)sums[3 * c + 2] + (long long)recv[c];
        long long carry = 0, gone = 0;
        bool done = false;
        for (uint32_t base = s0; base < s1 && !done; base += 1024) {
            const uint32_t i = base + (uint32_t)t;
            const bool in = i < s1;
            const uint32_t p = in ? pidx[i] : 0u;
            const long long wf = in ? (long long)km_fix((double)w[p], wscale) : 0ll;
            long long tot;
            const long long cum = carry + block_scan_incl(wf, wsum, &tot);
            if (t == 0) first = 1024;
            __syncthreads();
            if (in && W - cum <= limit) atomicMin(&first, t);
            __syncthreads();
            const int f = first;
            const bool mv = in && t <= f;
            if (mv) {
                const int32_t d = cand[p];
                lab[p] = d;
                atomicAdd(&recv[d], (unsigned long long)wf);
                ++moved_n;
            }
            if (f < 1024) {
                done = true;
                if (t == f) wsum[0] = cum;   // the weight handed over by this turn
                __syncthreads();
                gone = wsum[0];
            } else {
                carry += tot;
                gone = carry;
            }
            __syncthreads();
        }
        if (t == 0) cw[c] = (double)(W - gone) / wscale;
        __syncthreads();
    }
    if (moved_n) atomicAdd(ppm, moved_n);
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

// qsort_arg (src/utils.f90:239-298): the permutation sorting a[] in
// increasing order by the reference's partition scheme (middle pivot, scans
// with <= and >=, a meeting point stepped past), which fixes the order of
// equal values; used for the k cluster weights of the post-processing.
void qsort_arg_rec(std::vector<double> &a, std::vector<int32_t> &ord, int l, int r)
{
    const double pivot = a[(l + r) / 2];
    int i = l - 1, j = r + 1;
    for (;;) {
        do { --j; } while (!(a[j] <= pivot));
        do { ++i; } while (!(a[i] >= pivot));
        if (i < j) {
            std::swap(a[i], a[j]);
            std::swap(ord[i], ord[j]);
        } else {
            if (i == j) ++i;
            break;
        }
    }
    if (l < j) qsort_arg_rec(a, ord, l, j);
    if (i < r) qsort_arg_rec(a, ord, i, r);
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
    if (o->init != 0 && o->init != 1) return fail(ctx, TPCF_ERR_ARG, "init must be 0 or 1");
    if (!(o->weight_equalizer_scaling >= 0.0) || !std::isfinite(o->weight_equalizer_scaling))
        return fail(ctx, TPCF_ERR_ARG, "weight_equalizer_scaling must be finite and >= 0");
    if (o->equal_size != 0 && o->equal_size != 1) return fail(ctx, TPCF_ERR_ARG, "equal_size must be 0 or 1");
    if (o->equal_size && !(o->weight_tol >= 0.0)) return fail(ctx, TPCF_ERR_ARG, "weight_tol must be >= 0");
    if (flags & ~(uint32_t)(TPCF_DEVICE_INPUTS | TPCF_SPHERICAL)) return fail(ctx, TPCF_ERR_ARG, "unknown flag");
    const bool dev = (flags & TPCF_DEVICE_INPUTS) != 0, sph = (flags & TPCF_SPHERICAL) != 0;
    const double scaling = o->weight_equalizer_scaling;
    const bool eqs = o->equal_size != 0, eq = eqs || scaling > 0.0;
    TPCF_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;

    const unsigned nb = (unsigned)((n + kKmBlk - 1) / kKmBlk);
    const unsigned nbb = std::min<unsigned>(nb, 1024);
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t in_bytes = dev ? 0 : al(sizeof(T) * 3 * (size_t)n);
    const size_t lab_bytes = al(sizeof(int32_t) * n);
    const size_t small_bytes = al(sizeof(double) * (3 * k + 2 * k + k + 5 * nbb) + 64);
    // equal size: runner-up ratio and cluster per point, the post-processing's
    // keys, point indices and sort workspace
    size_t sort_tmp = 0;
    if (eqs) {
        TPCF_HIP(ctx, rocprim::radix_sort_pairs(nullptr, sort_tmp, (const double *)nullptr, (double *)nullptr,
                                                (const uint32_t *)nullptr, (uint32_t *)nullptr, (size_t)n, 0, 64, s));
        size_t t2 = 0;
        TPCF_HIP(ctx, rocprim::radix_sort_pairs(nullptr, t2, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                                (const uint32_t *)nullptr, (uint32_t *)nullptr, (size_t)n, 0, 11, s));
        sort_tmp = std::max(sort_tmp, t2);
    }
    const size_t eq_bytes = eqs ? al(sizeof(double) * n) * 2 + al(sizeof(int32_t) * n) + al(sizeof(uint32_t) * n) * 4 +
                                      al(sizeof(int32_t) * 2 * k + sizeof(uint32_t) * (k + 1) + 64) + al(sort_tmp)
                                : 0;
    char *wk;
    TPCF_TRY(ensure(ctx, tpcf_ctx::B_KMEANS, in_bytes + lab_bytes + small_bytes + eq_bytes, (void **)&wk));
    const T *dx = x, *dy = y, *dw = w;
    char *p = wk;
    if (!dev) {
        T *t = reinterpret_cast<T *>(p);
        TPCF_HIP(ctx, hipMemcpyAsync(t, x, sizeof(T) * n, hipMemcpyHostToDevice, s));
        TPCF_HIP(ctx, hipMemcpyAsync(t + n, y, sizeof(T) * n, hipMemcpyHostToDevice, s));
        TPCF_HIP(ctx, hipMemcpyAsync(t + 2 * n, w, sizeof(T) * n, hipMemcpyHostToDevice, s));
        dx = t; dy = t + n; dw = t + 2 * n;
        p += in_bytes;
    }
    int32_t *lab = reinterpret_cast<int32_t *>(p);
    p += lab_bytes;
    unsigned long long *sums = reinterpret_cast<unsigned long long *>(p);
    double *cen = reinterpret_cast<double *>(sums + 3 * k), *cw = cen + 2 * k, *bb = cw + k;
    unsigned long long *moves = reinterpret_cast<unsigned long long *>(bb + 5 * nbb);
    p += small_bytes;
    double *mem = nullptr, *mkey2 = nullptr;
    int32_t *cand = nullptr, *rank_d = nullptr, *order_d = nullptr;
    uint32_t *key = nullptr, *key2 = nullptr, *idx = nullptr, *idx2 = nullptr, *cnt = nullptr;
    void *tmp = nullptr;
    if (eqs) {
        mem = reinterpret_cast<double *>(p); p += al(sizeof(double) * n);
        mkey2 = reinterpret_cast<double *>(p); p += al(sizeof(double) * n);
        cand = reinterpret_cast<int32_t *>(p); p += al(sizeof(int32_t) * n);
        key = reinterpret_cast<uint32_t *>(p); p += al(sizeof(uint32_t) * n);
        key2 = reinterpret_cast<uint32_t *>(p); p += al(sizeof(uint32_t) * n);
        idx = reinterpret_cast<uint32_t *>(p); p += al(sizeof(uint32_t) * n);
        idx2 = reinterpret_cast<uint32_t *>(p); p += al(sizeof(uint32_t) * n);
        rank_d = reinterpret_cast<int32_t *>(p);
        order_d = rank_d + k;
        cnt = reinterpret_cast<uint32_t *>(order_d + k);
        p += al(sizeof(int32_t) * 2 * k + sizeof(uint32_t) * (k + 1) + 64);
        tmp = p;
    }

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
    // the fixed-point frame: |sum| <= n max|w| extent < 2^62 for every partial
    // sum of the coordinates; the weights in steps of a power of two with
    // n max|w| <= 2^52 (so single-precision weights sum exactly and every
    // weight sum converts to FP64 without rounding)
    KmFix fx;
    fx.x0 = xmin;
    fx.y0 = ymin;
    {
        const double wm = wmax > 0.0 ? wmax : 1.0, top = std::ldexp(1.0, 62);
        const double ex = xmax > xmin ? xmax - xmin : 1.0, ey = ymax > ymin ? ymax - ymin : 1.0;
        fx.scale[0] = top / ((double)n * wm * ex);
        fx.scale[1] = top / ((double)n * wm * ey);
        fx.scale[2] = std::ldexp(1.0, (int)std::floor(std::log2(std::ldexp(1.0, 52) / ((double)n * wm))));
    }
    double diam = std::sqrt((xmax - xmin) * (xmax - xmin) + (ymax - ymin) * (ymax - ymin)) * 10.0;
    // Voronoi form: compared as squared distances (chords on the sphere: 10
    // diameters exceed any chord); with the equaliser: distances, bounded by
    // 10 x the coordinate box diameter as the reference's
    const double init_best = eq ? diam : (sph ? 5.0 : diam * diam);
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
    if (eqs) {
        fill_f64(s, mem, n);   // ratio 0 and no runner-up before the second round
        hipLaunchKernelGGL(k_km_fill_i32, dim3(nb), dim3(kKmBlk), 0, s, n, cand, -1);
    }
    int it = 0;
    unsigned long long hm = 0;
    while (it <= o->max_iterations) {
        ++it;
        fill_u64(s, sums, 3 * k);
        fill_u64(s, moves, 1);
        double *mt = (eqs && it > 1) ? mem : nullptr;
        int32_t *ct = (eqs && it > 1) ? cand : nullptr;
        if (it > 1 && scaling > 0.0) {
            // the equaliser: one workgroup, the weights moving step by step
            if (sph)
                hipLaunchKernelGGL((k_km_assign_gs<T, true>), dim3(1), dim3(kKmGs), 0, s, n, k, dx, dy, dw, cen,
                                   init_best, fx, cw, scaling, lab, sums, moves, mt, ct);
            else
                hipLaunchKernelGGL((k_km_assign_gs<T, false>), dim3(1), dim3(kKmGs), 0, s, n, k, dx, dy, dw, cen,
                                   init_best, fx, cw, scaling, lab, sums, moves, mt, ct);
        } else {
#define TPCF_KM_ASSIGN(SP, EQ)                                                                                         \
    hipLaunchKernelGGL((k_km_assign<T, SP, EQ>), dim3(nb), dim3(kKmBlk), 0, s, n, k, dx, dy, dw, cen, init_best, fx, \
                       lab, sums, moves, (const double *)nullptr, mt, ct)
            if (sph) {
                if (eq) TPCF_KM_ASSIGN(true, true); else TPCF_KM_ASSIGN(true, false);
            } else {
                if (eq) TPCF_KM_ASSIGN(false, true); else TPCF_KM_ASSIGN(false, false);
            }
#undef TPCF_KM_ASSIGN
        }
        hipLaunchKernelGGL(k_km_centroids, dim3((k + 255) / 256), dim3(256), 0, s, k, sums, fx, cen, cw);
        TPCF_HIP(ctx, hipMemcpyAsync(&hm, moves, sizeof hm, hipMemcpyDeviceToHost, s));
        TPCF_HIP(ctx, hipStreamSynchronize(s));
        if (hm == 0 || (double)hm / (double)n < o->reassignment_tol) break;
    }
    TPCF_HIP(ctx, hipGetLastError());
    if (eqs) {
        // max weight deviation of the last round, and the post-processing
        // when it exceeds weight_tol (src/kmeans.f90:140-151)
        std::vector<unsigned long long> hs(3 * k);
        TPCF_HIP(ctx, hipMemcpyAsync(hs.data(), sums, sizeof(unsigned long long) * 3 * k, hipMemcpyDeviceToHost, s));
        TPCF_HIP(ctx, hipStreamSynchronize(s));
        long long tot = 0;
        std::vector<double> hw(k);
        for (int j = 0; j < k; ++j) {
            tot += (long long)hs[3 * j + 2];
            hw[j] = (double)(long long)hs[3 * j + 2] / fx.scale[2];
        }
        const double mean_w = (double)tot / fx.scale[2] / (double)k;
        const double dev_w = (*std::max_element(hw.begin(), hw.end()) - *std::min_element(hw.begin(), hw.end())) / mean_w;
        if (dev_w > o->weight_tol) {
            std::vector<int32_t> ord(k), rank(k);
            for (int j = 0; j < k; ++j) ord[j] = j;
            std::vector<double> wk2 = hw;
            qsort_arg_rec(wk2, ord, 0, k - 1);
            std::reverse(ord.begin(), ord.end());   // reverse = .true.: decreasing weight
            for (int r = 0; r < k; ++r) rank[ord[r]] = r;
            const double lim = mean_w * (1.0 + o->weight_tol) * fx.scale[2];
            const long long limit = lim >= 9.2e18 ? LLONG_MAX : (lim <= -9.2e18 ? LLONG_MIN : (long long)std::floor(lim));
            TPCF_HIP(ctx, hipMemcpyAsync(rank_d, rank.data(), sizeof(int32_t) * k, hipMemcpyHostToDevice, s));
            TPCF_HIP(ctx, hipMemcpyAsync(order_d, ord.data(), sizeof(int32_t) * k, hipMemcpyHostToDevice, s));
            TPCF_HIP(ctx, hipMemsetAsync(cnt, 0, sizeof(uint32_t) * (k + 1), s));
            hipLaunchKernelGGL(k_km_pp_keys, dim3(nb), dim3(kKmBlk), 0, s, n, k, lab, cand, rank_d, key, idx, cnt);
            // (cluster turn, ratio, index) order: a stable sort on the ratio, then on the turn
            size_t tb = sort_tmp;
            TPCF_HIP(ctx, rocprim::radix_sort_pairs(tmp, tb, (const double *)mem, mkey2, (const uint32_t *)idx, idx2,
                                                    (size_t)n, 0, 64, s));
            hipLaunchKernelGGL(k_km_gather_u32, dim3(nb), dim3(kKmBlk), 0, s, n, idx2, key, key2);
            tb = sort_tmp;
            TPCF_HIP(ctx, rocprim::radix_sort_pairs(tmp, tb, (const uint32_t *)key2, key, (const uint32_t *)idx2, idx,
                                                    (size_t)n, 0, 11, s));
            fill_u64(s, moves, 1);
            hipLaunchKernelGGL(k_km_pp_seq<T>, dim3(1), dim3(1024), 0, s, k, order_d, cnt, idx, lab, cand, dw,
                               fx.scale[2], sums, limit, cw, moves);
            TPCF_HIP(ctx, hipGetLastError());
        }
    }
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
