// sphere.hip -- spherical coordinates: the reference's spherical_coords =
// .true. branches of tangential_shear_tpcf / scalar_kappa_tpcf
// (src/tpcf.f90:313-320, :337-345), with distance_sphere
// (src/utils.f90:512-528), angles_sphere (src/utils.f90:618-639) and the
// spherical kd-tree search (src/kdtree.f90:377-590).  Positions are
// longitude x and latitude y in radians; bin edges are angles in radians.
//
// Geometry.  Every object becomes a unit vector u = (cos y cos x, cos y sin x,
// sin y) in FP64, expressed in a basis (c, e1, e2) whose first axis c is the
// mean direction of the lenses: p = u.c, q = u.e1, s = u.e2.
//   * The separation angle d and the chord |u_A - u_B| = 2 sin(d/2) grow
//     together on [0, pi], so bin k of the angle is the chord range
//     [2 sin(e_k/2), 2 sin(e_k+1/2)): the FP64 path bins squared chords against
//     squared chord edges and accumulates d = 2 asin(chord/2).
//   * The orthographic projection (q, s) never lengthens a distance, so the
//     flat-sky neighbour search (cells, segments, chunks) run on (q, s) with the
//     radius 2 sin(x_max/2) finds every pair; the pair kernel measures the 3-D
//     chord and bounds chunk-segment distances with the p range as well.
//   * angles_sphere's cos(theta) and sin(theta) are the components of u_B
//     along the east and north unit vectors at the lens, divided by sin d;
//     with E, N those components of u_B - u_A (the same numbers, since both
//     vectors are orthogonal to u_A), E^2 + N^2 = sin^2 d and
//     cos 2theta = (E^2 - N^2)/(E^2 + N^2), sin 2theta = 2EN/(E^2 + N^2).
//     The frames are those of the input coordinates, rotated with the basis.
// Everything runs in FP64: FP32 unit vectors hold 6e-8 absolute error, 4e-4
// of the chord of a 0.5' separation.  On gfx950 an FP64 FMA issues at the
// rate of an FP32 FMA.
#include <algorithm>
#include <cmath>
#include <vector>

#include "internal.hpp"

namespace tpcf {

namespace {

constexpr int kSphBlk = 256;

// Per-block sums of the unit vectors (the lenses' mean direction).
template <typename T>
__global__ __launch_bounds__(kSphBlk) void k_sph_sum(int64_t n, const T *__restrict__ lon, const T *__restrict__ lat,
                                                     double *__restrict__ part)
{
    __shared__ double sm[3][kSphBlk];
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)kSphBlk + threadIdx.x; i < n; i += (int64_t)gridDim.x * kSphBlk) {
        double sx, cx, sy, cy;
        sincos((double)lon[i], &sx, &cx);
        sincos((double)lat[i], &sy, &cy);
        a0 += cy * cx;
        a1 += cy * sx;
        a2 += sy;
    }
    sm[0][threadIdx.x] = a0;
    sm[1][threadIdx.x] = a1;
    sm[2][threadIdx.x] = a2;
    __syncthreads();
    for (int st = kSphBlk / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st)
            for (int k = 0; k < 3; ++k) sm[k][threadIdx.x] += sm[k][threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; ++k) part[blockIdx.x * 3 + k] = sm[k][0];
}

struct Basis {
    double c[3], e1[3], e2[3];
};

// Rotated unit vectors (p, q, s) and, for lenses, the rotated east and north
// unit vectors [n][6].
template <typename T>
__global__ __launch_bounds__(kSphBlk) void k_sph_coords(int64_t n, const T *__restrict__ lon,
                                                        const T *__restrict__ lat, const Basis B,
                                                        double *__restrict__ q, double *__restrict__ s,
                                                        double *__restrict__ p, double *__restrict__ frame)
{
    const int64_t i = blockIdx.x * (int64_t)kSphBlk + threadIdx.x;
    if (i >= n) return;
    double sx, cx, sy, cy;
    sincos((double)lon[i], &sx, &cx);
    sincos((double)lat[i], &sy, &cy);
    const double u[3] = {cy * cx, cy * sx, sy};
    auto dot = [](const double *a, const double *b) { return fma(a[0], b[0], fma(a[1], b[1], a[2] * b[2])); };
    p[i] = dot(u, B.c);
    q[i] = dot(u, B.e1);
    s[i] = dot(u, B.e2);
    if (frame) {
        const double east[3] = {-sx, cx, 0.0};
        const double north[3] = {-sy * cx, -sy * sx, cy};
        double *f = frame + 6 * i;
        f[0] = dot(east, B.c);
        f[1] = dot(east, B.e1);
        f[2] = dot(east, B.e2);
        f[3] = dot(north, B.c);
        f[4] = dot(north, B.e1);
        f[5] = dot(north, B.e2);
    }
}

template <typename T>
__global__ __launch_bounds__(kSphBlk) void k_sph_f64(int64_t n, const T *__restrict__ in, double *__restrict__ out)
{
    const int64_t i = blockIdx.x * (int64_t)kSphBlk + threadIdx.x;
    if (i < n) out[i] = (double)in[i];
}

// Sorted lenses: p and frames.
__global__ __launch_bounds__(kSphBlk) void k_sph_lens(int64_t n, const uint32_t *__restrict__ lidx,
                                                      const double *__restrict__ pa, const double *__restrict__ fa,
                                                      double *__restrict__ lp, double *__restrict__ lf)
{
    const int64_t i = blockIdx.x * (int64_t)kSphBlk + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = lidx[i];
    lp[i] = pa[j];
    for (int k = 0; k < 6; ++k) lf[6 * i + k] = fa[6 * (size_t)j + k];
}

// Sorted sources: p into record slot s1w2.
__global__ __launch_bounds__(kSphBlk) void k_sph_src(int64_t n, const uint32_t *__restrict__ sidx,
                                                     const double *__restrict__ pb, SrcRec<double> *__restrict__ rec)
{
    const int64_t i = blockIdx.x * (int64_t)kSphBlk + threadIdx.x;
    if (i < n) rec[i].s1w2 = pb[sidx[i]];
}

// (min, max) of p over each segment's records.
__global__ __launch_bounds__(kSphBlk) void k_sph_segp(int64_t n, const Seg<double> *__restrict__ segs,
                                                      const SrcRec<double> *__restrict__ rec,
                                                      double2 *__restrict__ segp)
{
    const int64_t i = blockIdx.x * (int64_t)kSphBlk + threadIdx.x;
    if (i >= n) return;
    const uint32_t s0 = segs[i].start, e = s0 + segs[i].count;
    double lo = INFINITY, hi = -INFINITY;
    for (uint32_t k = s0; k < e; ++k) {
        lo = fmin(lo, rec[k].s1w2);
        hi = fmax(hi, rec[k].s1w2);
    }
    segp[i] = make_double2(lo, hi);
}

// (min, max) of p over each chunk's lenses.
__global__ __launch_bounds__(kSphBlk) void k_sph_chp(int64_t n, const Chunk<double> *__restrict__ chunks,
                                                     const double *__restrict__ lp, double2 *__restrict__ chp)
{
    const int64_t i = blockIdx.x * (int64_t)kSphBlk + threadIdx.x;
    if (i >= n) return;
    const uint32_t s0 = chunks[i].start, e = s0 + chunks[i].count;
    double lo = INFINITY, hi = -INFINITY;
    for (uint32_t k = s0; k < e; ++k) {
        lo = fmin(lo, lp[k]);
        hi = fmax(hi, lp[k]);
    }
    chp[i] = make_double2(lo, hi);
}

// Frames for a pass with the catalogues' roles swapped: (u x P, P) of each
// object from its rotated unit vector (p, q, s); the basis is right-handed
// (c x e1 = e2), so the cross product is taken in rotated coordinates.
__global__ __launch_bounds__(kSphBlk) void k_sph_rev_frame(int64_t n, const double *__restrict__ p,
                                                           const double *__restrict__ q,
                                                           const double *__restrict__ s, const double P0,
                                                           const double P1, const double P2,
                                                           double *__restrict__ frame)
{
    const int64_t i = blockIdx.x * (int64_t)kSphBlk + threadIdx.x;
    if (i >= n) return;
    const double u0 = p[i], u1 = q[i], u2 = s[i];
    double *f = frame + 6 * i;
    f[0] = fma(u1, P2, -(u2 * P1));
    f[1] = fma(u2, P0, -(u0 * P2));
    f[2] = fma(u0, P1, -(u1 * P0));
    f[3] = P0;
    f[4] = P1;
    f[5] = P2;
}

inline unsigned sph_blocks(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + kSphBlk - 1) / kSphBlk); }

void cross(const double *a, const double *b, double *o)
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

void normalise(double *v)
{
    const double n = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    for (int k = 0; k < 3; ++k) v[k] /= n;
}

// Orthonormal basis with first axis along the mean direction of (lon, lat).
template <typename T>
int sph_basis(tpcf_ctx *ctx, int64_t n, const T *lon, const T *lat, Basis &B)
{
    double c[3] = {1.0, 0.0, 0.0};
    if (n > 0) {
        const unsigned nb = std::min<unsigned>(1024, sph_blocks(n));
        double *part;
        TPCF_TRY(ensure(ctx, tpcf_ctx::B_SPH_PART, sizeof(double) * 3 * nb, (void **)&part));
        hipLaunchKernelGGL(k_sph_sum<T>, dim3(nb), dim3(kSphBlk), 0, ctx->stream, n, lon, lat, part);
        std::vector<double> h(3 * nb);
        TPCF_HIP(ctx, hipMemcpyAsync(h.data(), part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, ctx->stream));
        TPCF_HIP(ctx, hipStreamSynchronize(ctx->stream));
        double m[3] = {0.0, 0.0, 0.0};
        for (unsigned b = 0; b < nb; ++b)
            for (int k = 0; k < 3; ++k) m[k] += h[3 * b + k];
        const double len = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
        // any axis gives the same pairs; a mean near zero (a symmetric all-sky
        // catalogue) keeps the default
        if (std::isfinite(len) && len > 1e-6 * (double)n)
            for (int k = 0; k < 3; ++k) c[k] = m[k] / len;
    }
    // e1: the coordinate axis least aligned with c, made orthogonal to it
    int kmin = 0;
    for (int k = 1; k < 3; ++k)
        if (std::fabs(c[k]) < std::fabs(c[kmin])) kmin = k;
    double a[3] = {0.0, 0.0, 0.0};
    a[kmin] = 1.0;
    double e2[3], e1[3];
    cross(c, a, e2);
    normalise(e2);
    cross(e2, c, e1);
    normalise(e1);
    for (int k = 0; k < 3; ++k) {
        B.c[k] = c[k];
        B.e1[k] = e1[k];
        B.e2[k] = e2[k];
    }
    return TPCF_OK;
}

}  // namespace

int sph_chord_bins(tpcf_ctx *ctx, const tpcf_bins *bins, tpcf_bins &kb)
{
    const int n_bin = bins->n_bin;
    if (!(bins->edges[0] >= 0.0))
        return fail(ctx, TPCF_ERR_BINS, "spherical bins must be non-negative angles in radians");
    // chord edges 2 sin(e/2) for e in [0, pi]; angles past pi hold no pairs
    kb = *bins;
    const double pi = 3.14159265358979323846;
    for (int i = 0; i <= n_bin; ++i) kb.edges[i] = 2.0 * std::sin(0.5 * std::min(bins->edges[i], pi));
    for (int i = 1; i <= n_bin; ++i)   // keep them increasing past pi (those bins stay empty)
        if (!(kb.edges[i] > kb.edges[i - 1])) kb.edges[i] = std::nextafter(kb.edges[i - 1], INFINITY);
    kb.x_min = (float)kb.edges[0];
    kb.x_max = (float)kb.edges[n_bin];
    return TPCF_OK;
}

void launch_to_f64(hipStream_t s, const float *in, double *out, int64_t n)
{
    if (n > 0) hipLaunchKernelGGL(k_sph_f64<float>, dim3(sph_blocks(n)), dim3(kSphBlk), 0, s, n, in, out);
}

void sph_records_p(hipStream_t s, const Prepared<double> &p, const double *pb)
{
    if (p.n_src > 0)
        hipLaunchKernelGGL(k_sph_src, dim3(sph_blocks(p.n_src)), dim3(kSphBlk), 0, s, p.n_src, p.sidx, pb,
                           const_cast<SrcRec<double> *>(p.srec));
}

template <typename T>
int sph_setup(tpcf_ctx *ctx, const CallArgs<T> &a, const tpcf_bins *kb, CallArgs<double> &d,
              std::vector<F64Copy> *copies)
{
    hipStream_t s = ctx->stream;
    const int64_t n_a = a.n_a, n_b = a.n_b;
    const bool shear = (a.kind == SHEAR);
    const tpcf_bins *bins = a.bins;

    Basis B;
    if (n_a > 0) TPCF_TRY(sph_basis<T>(ctx, n_a, a.xa, a.ya, B));
    else TPCF_TRY(sph_basis<T>(ctx, n_b, a.xb, a.yb, B));

    // FP64 work arrays: lens q, s, p, frames [n_a][6], (z, val, w); source q, s, p, (z, s1, s2, w, m)
    const int64_t na1 = std::max<int64_t>(n_a, 1), nb1 = std::max<int64_t>(n_b, 1);
    const bool conv = (sizeof(T) == 4);
    const int64_t need = na1 * (9 + (conv ? 3 : 0)) + nb1 * (3 + (conv ? 5 : 0));
    double *wk;
    TPCF_TRY(ensure(ctx, tpcf_ctx::B_SPH_IN, sizeof(double) * need, (void **)&wk));
    double *qa = wk, *sa = qa + na1, *pa = sa + na1, *fa = pa + na1, *qb = fa + 6 * na1, *sbv = qb + nb1,
           *pbv = sbv + nb1, *rest = pbv + nb1;
    if (n_a > 0)
        hipLaunchKernelGGL(k_sph_coords<T>, dim3(sph_blocks(n_a)), dim3(kSphBlk), 0, s, n_a, a.xa, a.ya, B, qa, sa,
                           pa, fa);
    if (n_b > 0)
        hipLaunchKernelGGL(k_sph_coords<T>, dim3(sph_blocks(n_b)), dim3(kSphBlk), 0, s, n_b, a.xb, a.yb, B, qb, sbv,
                           pbv, (double *)nullptr);
    auto f64 = [&](const T *in, int64_t n, double *&slot) -> const double * {
        if (!in) return nullptr;
        if constexpr (sizeof(T) == 8) {
            return reinterpret_cast<const double *>(in);
        } else {
            double *o = slot;
            slot += std::max<int64_t>(n, 1);
            if (n > 0) hipLaunchKernelGGL(k_sph_f64<T>, dim3(sph_blocks(n)), dim3(kSphBlk), 0, s, n, in, o);
            if (copies) copies->push_back(F64Copy{reinterpret_cast<const float *>(in), o, n});
            return o;
        }
    };
    d = CallArgs<double>{};
    d.kind = a.kind;
    d.n_a = n_a;
    d.n_b = n_b;
    d.n_pa = a.n_pa;
    d.n_pb = a.n_pb;
    d.patch_a = a.patch_a;
    d.patch_b = a.patch_b;
    d.bins = bins;
    d.kbins = kb;
    d.marks = a.marks;
    d.verbose = a.verbose;
    d.xa = qa;
    d.ya = sa;
    d.xb = qb;
    d.yb = sbv;
    double *slot = rest;
    d.za = shear ? f64(a.za, n_a, slot) : nullptr;
    d.vala = f64(a.vala, n_a, slot);
    d.wa = f64(a.wa, n_a, slot);
    d.zb = shear ? f64(a.zb, n_b, slot) : nullptr;
    d.s1b = f64(a.s1b, n_b, slot);
    d.s2b = f64(a.s2b, n_b, slot);
    d.wb = f64(a.wb, n_b, slot);
    d.mb = shear ? f64(a.mb, n_b, slot) : nullptr;
    d.sph_pa = pa;
    d.sph_pb = pbv;
    d.sph_frame = fa;
    d.sph_pole[0] = B.c[2];
    d.sph_pole[1] = B.e1[2];
    d.sph_pole[2] = B.e2[2];
    TPCF_HIP(ctx, hipGetLastError());
    return TPCF_OK;
}

int sph_rev_frames(tpcf_ctx *ctx, const CallArgs<double> &d, const double **frames)
{
    double *f;
    TPCF_TRY(ensure(ctx, tpcf_ctx::B_SPH_RF, sizeof(double) * 6 * std::max<int64_t>(d.n_b, 1), (void **)&f));
    if (d.n_b > 0)
        hipLaunchKernelGGL(k_sph_rev_frame, dim3(sph_blocks(d.n_b)), dim3(kSphBlk), 0, ctx->stream, d.n_b, d.sph_pb,
                           d.xb, d.yb, d.sph_pole[0], d.sph_pole[1], d.sph_pole[2], f);
    TPCF_HIP(ctx, hipGetLastError());
    *frames = f;
    return TPCF_OK;
}

template <typename T>
int sph_correlate(tpcf_ctx *ctx, const CallArgs<T> &a, uint32_t flags, tpcf_out *out, double t0)
{
    tpcf_bins kb;
    TPCF_TRY(sph_chord_bins(ctx, a.bins, kb));
    CallArgs<double> d;
    TPCF_TRY(sph_setup<T>(ctx, a, &kb, d, nullptr));
    return run_call<double>(ctx, d, flags, out, t0, nullptr);
}

int sph_finish(tpcf_ctx *ctx, const CallArgs<double> &a, Prepared<double> &p)
{
    hipStream_t s = ctx->stream;
    double *lp, *lf;
    double2 *segp, *chp;
    TPCF_TRY(ensure(ctx, tpcf_ctx::B_SPH_LP, sizeof(double) * std::max<int64_t>(p.n_lens, 1), (void **)&lp));
    TPCF_TRY(ensure(ctx, tpcf_ctx::B_SPH_LF, sizeof(double) * 6 * std::max<int64_t>(p.n_lens, 1), (void **)&lf));
    TPCF_TRY(ensure(ctx, tpcf_ctx::B_SPH_SEGP, sizeof(double2) * std::max<int64_t>(p.n_segs, 1), (void **)&segp));
    TPCF_TRY(ensure(ctx, tpcf_ctx::B_SPH_CHP, sizeof(double2) * std::max<int64_t>(p.n_chunks, 1), (void **)&chp));
    if (p.n_lens > 0)
        hipLaunchKernelGGL(k_sph_lens, dim3(sph_blocks(p.n_lens)), dim3(kSphBlk), 0, s, p.n_lens, p.lidx, a.sph_pa,
                           a.sph_frame, lp, lf);
    if (p.n_src > 0)
        hipLaunchKernelGGL(k_sph_src, dim3(sph_blocks(p.n_src)), dim3(kSphBlk), 0, s, p.n_src, p.sidx, a.sph_pb,
                           const_cast<SrcRec<double> *>(p.srec));
    if (p.n_segs > 0)
        hipLaunchKernelGGL(k_sph_segp, dim3(sph_blocks(p.n_segs)), dim3(kSphBlk), 0, s, p.n_segs, p.segs, p.srec,
                           segp);
    if (p.n_chunks > 0)
        hipLaunchKernelGGL(k_sph_chp, dim3(sph_blocks(p.n_chunks)), dim3(kSphBlk), 0, s, p.n_chunks, p.chunks, lp,
                           chp);
    TPCF_HIP(ctx, hipGetLastError());
    p.sph_lp = lp;
    p.sph_lf = lf;
    p.sph_segp = segp;
    p.sph_chp = chp;
    p.sph_rev = a.sph_rev;
    return TPCF_OK;
}

template int sph_setup<float>(tpcf_ctx *, const CallArgs<float> &, const tpcf_bins *, CallArgs<double> &,
                               std::vector<F64Copy> *);
template int sph_setup<double>(tpcf_ctx *, const CallArgs<double> &, const tpcf_bins *, CallArgs<double> &,
                                std::vector<F64Copy> *);
template int sph_correlate<float>(tpcf_ctx *, const CallArgs<float> &, uint32_t, tpcf_out *, double);
template int sph_correlate<double>(tpcf_ctx *, const CallArgs<double> &, uint32_t, tpcf_out *, double);

}  // namespace tpcf
