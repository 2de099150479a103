// adjoint.hip -- reverse mode of the per-bin outputs: given the gradient of a
// scalar loss with respect to g_t / g_x (shear) or xi_E / xi_B (kappa) per bin,
// the gradient with respect to every object's value and weight.  The
// reference has no derivative; the forward quantities are those of
// tangential_shear_tpcf (src/tpcf.f90:340-352, :369-376) and scalar_kappa_tpcf
// (src/tpcf.f90:512-519, :536-543).
//
// With a_i = val_i w_i and per bin b (sums over the pairs (i, j) in bin b):
//   shear  G_b = -sum a_i w_j (e1_j c_ij + e2_j s_ij),  X_b = -sum a_i w_j (-e1_j s_ij + e2_j c_ij),
//          K_b = sum w_i w_j (1 + m_j),  g_t = G/K, g_x = X/K;
//   kappa  E_b = sum a_i w_j kE_j,  B_b = sum a_i w_j kB_j,  W_b = sum w_i w_j,  xi = E/W, B/W.
// With upstream gradients lt_b, lx_b, alpha_b = lt_b / K_b, alpha'_b = lx_b / K_b and
// beta_b = (lt_b g_t,b + lx_b g_x,b) / K_b, the gradient is a per-object
// contraction over bins of per-object sums over its partners:
//   lens i:    P_ib = -sum_j w_j (e1_j c + e2_j s), Q_ib = -sum_j w_j (-e1_j s + e2_j c), kap_ib = sum_j w_j (1 + m_j)
//              dval_i = w_i sum_b (alpha P + alpha' Q),  dw_i = val_i sum_b (alpha P + alpha' Q) - sum_b beta kap
//   source j:  C_jb = -sum_i a_i c_ij, S'_jb = sum_i a_i s_ij, Om_jb = sum_i w_i
//              de1_j = w_j sum_b (alpha C + alpha' S'),  de2_j = w_j sum_b (-alpha S' + alpha' C),
//              dw_j = sum_b (alpha (e1 C - e2 S') + alpha' (e1 S' + e2 C)) - (1 + m_j) sum_b beta Om,
//              dm_j = -w_j sum_b beta Om
// and for kappa (S^E_ib = sum_j w_j kE_j, S^B_ib, Om_ib = sum_j w_j; A_jb = sum_i a_i, Om_jb = sum_i w_i):
//              dval_i = w_i sum_b (alpha S^E + alpha' S^B),  dw_i = val_i sum_b (...) - sum_b beta Om_ib,
//              dkE_j = w_j sum_b alpha A,  dkB_j = w_j sum_b alpha' A,  dw_j = sum_b (alpha kE_j + alpha' kB_j) A - sum_b beta Om_jb.
// The per-object sums are the pair kernel's marks: once with unit lens values
// and weights (lens side; the forward per-bin sums G, X, K are contractions of
// these marks with a_i and w_i, so no forward pass of its own), once with the
// roles of the catalogues swapped
// (source side: the sources become the lanes, the lenses are streamed with
// their a_i as the first component; the z test is reversed by storing -z).
// Lenses with w = 0 (skipped by the reference) get the gradient of the
// smooth extension; objects masked by z < 0 get zero.
#include <cmath>
#include <vector>

#include "internal.hpp"

namespace tpcf {

void fill_f64(hipStream_t s, double *p, int64_t n);

namespace {

struct Coef {
    double at[kGroupBins], ax[kGroupBins], be[kGroupBins];
};

template <typename T>
__global__ void k_fill_value(T *p, int64_t n, T v)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// The forward sums from the lens-side marks (unit lens values and weights):
// G_b = sum_i a_i P_ib, X_b = sum_i a_i Q_ib, K_b = sum_i w_i kap_ib (kappa:
// E, B and W with S^E, S^B, Om) -- the per-bin sums of the forward pass
// without a pass of their own.  blockIdx.y = bin; out[b][3] by atomics.
template <typename T>
__global__ __launch_bounds__(256) void k_adj_fwd(int64_t n_lens, int32_t n_bin, const uint32_t *__restrict__ lidx,
                                                 const double *__restrict__ mk, const T *__restrict__ val,
                                                 const T *__restrict__ w, int kind, double *__restrict__ out)
{
    __shared__ double sm[3][256];
    const int b = blockIdx.y;
    double g = 0.0, x = 0.0, k = 0.0;
    for (int64_t li = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; li < n_lens;
         li += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t o = lidx[li];
        const double wi = (double)w[o], ai = (double)val[o] * wi;
        g = fma(ai, mk[mark_at(n_lens, b, 0, li)], g);
        x = fma(ai, mk[mark_at(n_lens, b, 1, li)], x);
        k = fma(wi, mk[mark_at(n_lens, b, kind == SHEAR ? 2 : 3, li)], k);
    }
    sm[0][threadIdx.x] = g;
    sm[1][threadIdx.x] = x;
    sm[2][threadIdx.x] = k;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st)
            for (int q = 0; q < 3; ++q) sm[q][threadIdx.x] += sm[q][threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 3; ++q) unsafeAtomicAdd(&out[3 * b + q], sm[q][0]);
}

// Lens side: the marks of sorted lens li in bin b (mark_at planes) are
// (P or S^E, Q or S^B, kap, Om).
template <typename T>
__global__ void k_adj_lens(int64_t n_lens, int32_t n_bin, const uint32_t *__restrict__ lidx,
                           const double *__restrict__ mk, const T *__restrict__ val, const T *__restrict__ w,
                           const Coef c, double *__restrict__ dval, double *__restrict__ dw)
{
    const int64_t li = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (li >= n_lens) return;
    const uint32_t o = lidx[li];
    double s = 0.0, sk = 0.0;
    for (int b = 0; b < n_bin; ++b) {
        s += c.at[b] * mk[mark_at(n_lens, b, 0, li)] + c.ax[b] * mk[mark_at(n_lens, b, 1, li)];
        sk += c.be[b] * mk[mark_at(n_lens, b, 2, li)];
    }
    if (dval) dval[o] = (double)w[o] * s;
    if (dw) dw[o] = (double)val[o] * s - sk;
}

// Source side (roles swapped): the marks of sorted source li in bin b are
// (C or A, S', ., Om) of the
// sorted source li.
template <typename T, int KIND>
__global__ void k_adj_src(int64_t n_src, int32_t n_bin, const uint32_t *__restrict__ lidx,
                          const double *__restrict__ mk, const T *__restrict__ s1, const T *__restrict__ s2,
                          const T *__restrict__ w, const T *__restrict__ m, const Coef c, double *__restrict__ d1,
                          double *__restrict__ d2, double *__restrict__ dw, double *__restrict__ dm)
{
    const int64_t li = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (li >= n_src) return;
    const uint32_t o = lidx[li];
    const double e1 = (double)s1[o], e2 = (double)s2[o], wj = (double)w[o];
    const double opm = (KIND == SHEAR && m) ? 1.0 + (double)m[o] : 1.0;
    double a1 = 0.0, a2 = 0.0, aw = 0.0, bo = 0.0;
    for (int b = 0; b < n_bin; ++b) {
        const double P = mk[mark_at(n_src, b, 0, li)], Q = mk[mark_at(n_src, b, 1, li)],
                     Om = mk[mark_at(n_src, b, 3, li)];
        if (KIND == SHEAR) {
            a1 += c.at[b] * P + c.ax[b] * Q;
            a2 += -c.at[b] * Q + c.ax[b] * P;
            aw += c.at[b] * (e1 * P - e2 * Q) + c.ax[b] * (e1 * Q + e2 * P);
        } else {
            a1 += c.at[b] * P;
            a2 += c.ax[b] * P;
            aw += (c.at[b] * e1 + c.ax[b] * e2) * P;
        }
        bo += c.be[b] * Om;
    }
    if (d1) d1[o] = wj * a1;
    if (d2) d2[o] = wj * a2;
    if (dw) dw[o] = aw - opm * bo;
    if (dm) dm[o] = -wj * bo;
}

template <typename T>
int upload_in(tpcf_ctx *ctx, int which, const T *src, int64_t n, const T **dst)
{
    if (!src) {
        *dst = nullptr;
        return TPCF_OK;
    }
    void *p;
    TPCF_TRY(ensure(ctx, which, sizeof(T) * std::max<int64_t>(n, 1), &p));
    if (n > 0) TPCF_HIP(ctx, hipMemcpyAsync(p, src, sizeof(T) * n, hipMemcpyHostToDevice, ctx->stream));
    *dst = static_cast<const T *>(p);
    return TPCF_OK;
}

// The lens-side pass and what the rest of the adjoint needs from it: the
// marks of the sorted lenses (P or S^E, Q or S^B, kap, Om per bin), the lens
// order and the forward per-bin sums G, X, K (kappa: E, B, W).
struct LensSide {
    const double *mk = nullptr;
    const uint32_t *lidx = nullptr;
    int64_t n_lens = 0;
    std::vector<double> fwd;   // [n_bin][3]
};

// Device arrays of n ones and n zeros (the unit values and weights of the
// passes), filled on the context's stream.
template <typename T>
int unit_arrays(tpcf_ctx *ctx, int64_t n, T **ones, T **zeros)
{
    const int64_t n_max = std::max<int64_t>(n, 1);
    TPCF_TRY(ensure(ctx, tpcf_ctx::B_ONES, sizeof(T) * n_max, (void **)ones));
    TPCF_TRY(ensure(ctx, tpcf_ctx::B_ZEROS, sizeof(T) * n_max, (void **)zeros));
    const unsigned fb = (unsigned)std::min<int64_t>(4096, (n_max + 255) / 256);
    hipLaunchKernelGGL(k_fill_value<T>, dim3(fb), dim3(256), 0, ctx->stream, *ones, n_max, (T)1);
    hipLaunchKernelGGL(k_fill_value<T>, dim3(fb), dim3(256), 0, ctx->stream, *zeros, n_max, (T)0);
    TPCF_HIP(ctx, hipGetLastError());
    return TPCF_OK;
}

// The geometry the passes run on.  Flat sky: the inputs themselves.  Sphere:
// the FP64 projected positions and values of sph_setup (one basis for both
// passes), the chord bins, and the sources' frames for the swapped pass.  The
// per-object contractions read the caller's values in both cases.
struct SphGeo {
    tpcf_bins kb{};
    CallArgs<double> d;
    const double *rev_frames = nullptr;
};

template <typename T>
int sph_geo(tpcf_ctx *ctx, const CallArgs<T> &base, SphGeo &g)
{
    TPCF_TRY(sph_chord_bins(ctx, base.bins, g.kb));
    TPCF_TRY(sph_setup<T>(ctx, base, &g.kb, g.d, nullptr));
    TPCF_TRY(sph_rev_frames(ctx, g.d, &g.rev_frames));
    return TPCF_OK;
}

// Lens side: unit lens values and weights, real sources.  The marks stay in
// the context's marks buffer and the lens order in its preparation buffers
// (both valid until the next pass on this context); the forward sums are
// contracted from the marks with a_i and w_i.  geo: the pass's geometry (U = T
// on the flat sky, FP64 on the sphere).
template <typename T, typename U>
int lens_pass(tpcf_ctx *ctx, const CallArgs<T> &base, const CallArgs<U> &geo, uint32_t run_flags, double t0,
              LensSide &ls)
{
    hipStream_t s = ctx->stream;
    const int n_bin = base.bins->n_bin;
    ls.fwd.assign(3 * (size_t)n_bin, 0.0);
    ls.n_lens = 0;
    if (base.n_a == 0 || base.n_b == 0) return TPCF_OK;
    U *ones, *zeros;
    TPCF_TRY(unit_arrays<U>(ctx, std::max(base.n_a, base.n_b), &ones, &zeros));
    CallArgs<U> l = geo;
    l.vala = ones;
    l.wa = ones;
    l.marks = true;
    l.marks_device = true;
    tpcf_out none;
    std::memset(&none, 0, sizeof none);
    none.real_bytes = 8;
    Prepared<U> p;
    TPCF_TRY(run_call<U>(ctx, l, run_flags, &none, t0, &p));
    ls.mk = static_cast<const double *>(ctx->buf[tpcf_ctx::B_MARKS].p);
    ls.lidx = p.lidx;
    ls.n_lens = p.n_lens;
    if (p.n_lens > 0) {
        double *fs;
        TPCF_TRY(ensure(ctx, tpcf_ctx::B_TMP, sizeof(double) * 3 * n_bin, (void **)&fs));
        fill_f64(s, fs, 3 * n_bin);
        const unsigned gx = (unsigned)std::min<int64_t>(1024, (p.n_lens + 255) / 256);
        hipLaunchKernelGGL(k_adj_fwd<T>, dim3(gx, (unsigned)n_bin), dim3(256), 0, s, p.n_lens, n_bin, p.lidx, ls.mk,
                           base.vala, base.wa, base.kind, fs);
        TPCF_HIP(ctx, hipGetLastError());
        TPCF_HIP(ctx, hipMemcpyAsync(ls.fwd.data(), fs, sizeof(double) * ls.fwd.size(), hipMemcpyDeviceToHost, s));
        TPCF_HIP(ctx, hipStreamSynchronize(s));
    }
    return TPCF_OK;
}

// Per-bin coefficients of the contraction: alpha = lbar_1 / K, alpha' =
// lbar_2 / K, beta = (lbar_1 g_1 + lbar_2 g_2) / K, zero where K <= 0 (the
// output is 0 there, with no dependence).
Coef coef_of(const std::vector<double> &fwd, int n_bin, const double *gbar1, const double *gbar2)
{
    Coef c;
    std::memset(&c, 0, sizeof c);
    for (int b = 0; b < n_bin; ++b) {
        const double G = fwd[3 * b], X = fwd[3 * b + 1], K = fwd[3 * b + 2];
        if (!(K > 0.0)) continue;
        const double g1 = G / K, g2 = X / K;
        c.at[b] = gbar1[b] / K;
        c.ax[b] = gbar2[b] / K;
        c.be[b] = (gbar1[b] * g1 + gbar2[b] * g2) / K;
    }
    return c;
}

template <typename T>
int lens_grad(tpcf_ctx *ctx, const LensSide &ls, const CallArgs<T> &base, const Coef &c, double *dvA, double *dwA)
{
    if (ls.n_lens == 0 || !(dvA || dwA)) return TPCF_OK;
    hipLaunchKernelGGL(k_adj_lens<T>, dim3((unsigned)((ls.n_lens + 255) / 256)), dim3(256), 0, ctx->stream, ls.n_lens,
                       base.bins->n_bin, ls.lidx, ls.mk, base.vala, base.wa, c, dvA, dwA);
    TPCF_HIP(ctx, hipGetLastError());
    return TPCF_OK;
}

// Source side: the sources as lanes, the lenses streamed with (a_i, 0).  On
// the sphere the shear angle stays the one at the lens, now the streamed
// object (CallArgs::sph_rev, with the sources' frames rev_frames).
template <typename T, typename U>
int src_pass(tpcf_ctx *ctx, const CallArgs<T> &base, const CallArgs<U> &geo, const double *rev_frames,
             uint32_t run_flags, double t0, const Coef &c, double *d1B, double *d2B, double *dwB, double *dmB)
{
    if (base.n_a == 0 || base.n_b == 0 || !(d1B || d2B || dwB || dmB)) return TPCF_OK;
    hipStream_t s = ctx->stream;
    const bool shear = base.kind == SHEAR;
    U *ones, *zeros;
    TPCF_TRY(unit_arrays<U>(ctx, std::max(base.n_a, base.n_b), &ones, &zeros));
    CallArgs<U> r = geo;
    r.n_a = geo.n_b;
    r.xa = geo.xb; r.ya = geo.yb; r.za = geo.zb;
    r.vala = ones;
    r.wa = ones;
    r.n_b = geo.n_a;
    r.xb = geo.xa; r.yb = geo.ya; r.zb = geo.za;
    r.s1b = geo.vala;
    r.s2b = zeros;
    r.wb = geo.wa;
    r.mb = shear ? zeros : nullptr;
    r.zsign = -1;   // the z test z_A > z_B of the original roles
    r.marks = true;
    r.marks_device = true;
    if (geo.sph_pa) {
        r.sph_pa = geo.sph_pb;
        r.sph_pb = geo.sph_pa;
        r.sph_frame = rev_frames;
        r.sph_rev = shear;
    }
    tpcf_out none;
    std::memset(&none, 0, sizeof none);
    none.real_bytes = 8;
    Prepared<U> p;
    TPCF_TRY(run_call<U>(ctx, r, run_flags, &none, t0, &p));
    const double *mk = static_cast<const double *>(ctx->buf[tpcf_ctx::B_MARKS].p);
    if (p.n_lens > 0) {
        const int n_bin = base.bins->n_bin;
        if (shear)
            hipLaunchKernelGGL((k_adj_src<T, SHEAR>), dim3((unsigned)((p.n_lens + 255) / 256)), dim3(256), 0, s,
                               p.n_lens, n_bin, p.lidx, mk, base.s1b, base.s2b, base.wb, base.mb, c, d1B, d2B, dwB,
                               dmB);
        else
            hipLaunchKernelGGL((k_adj_src<T, KAPPA>), dim3((unsigned)((p.n_lens + 255) / 256)), dim3(256), 0, s,
                               p.n_lens, n_bin, p.lidx, mk, base.s1b, base.s2b, base.wb, (const T *)nullptr, c, d1B,
                               d2B, dwB, (double *)nullptr);
        TPCF_HIP(ctx, hipGetLastError());
    }
    return TPCF_OK;
}

// Argument checks shared by the adjoint entry points.
template <typename T>
int adj_check(tpcf_ctx *ctx, int kind, int64_t n_a, const T *xa, const T *ya, const T *za, const T *vala,
              const T *wa, int64_t n_b, const T *xb, const T *yb, const T *zb, const T *s1b, const T *s2b,
              const T *wb, const T *mb, const tpcf_bins *bins, uint32_t flags)
{
    TPCF_TRY(check_bins(ctx, bins));
    if (bins->n_bin > kGroupBins)
        return fail(ctx, TPCF_ERR_UNSUPPORTED, "the adjoint takes at most " + std::to_string(kGroupBins) + " bins");
    if (flags & ~(uint32_t)(TPCF_DEVICE_INPUTS | TPCF_VERBOSE | TPCF_NO_KD_TREE | TPCF_SPHERICAL))
        return fail(ctx, TPCF_ERR_ARG, "unknown flag");
    if (n_a < 0 || n_b < 0 || n_a >= ((int64_t)1 << 31) || n_b >= ((int64_t)1 << 31))
        return fail(ctx, TPCF_ERR_ARG, "catalogue sizes must be in [0, 2^31)");
    const bool shear = (kind == SHEAR);
    if (n_a > 0 && (!xa || !ya || !vala || !wa || (shear && !za)))
        return fail(ctx, TPCF_ERR_ARG, "a required lens array is NULL");
    if (n_b > 0 && (!xb || !yb || !s1b || !s2b || !wb || (shear && (!zb || !mb))))
        return fail(ctx, TPCF_ERR_ARG, "a required source array is NULL");
    return TPCF_OK;
}

template <typename T>
CallArgs<T> adj_base(int kind, int64_t n_a, const T *xa, const T *ya, const T *za, const T *vala, const T *wa,
                     int64_t n_b, const T *xb, const T *yb, const T *zb, const T *s1b, const T *s2b, const T *wb,
                     const T *mb, const tpcf_bins *bins, uint32_t flags)
{
    CallArgs<T> base;
    base.kind = kind;
    base.n_a = n_a;
    base.n_b = n_b;
    base.n_pa = base.n_pb = 1;
    base.patch_a = base.patch_b = nullptr;
    base.bins = bins;
    base.marks = false;
    base.verbose = (flags & TPCF_VERBOSE) != 0;
    base.xa = xa; base.ya = ya; base.za = za; base.vala = vala; base.wa = wa;
    base.xb = xb; base.yb = yb; base.zb = zb; base.s1b = s1b; base.s2b = s2b; base.wb = wb; base.mb = mb;
    if (kind != SHEAR) { base.za = nullptr; base.zb = nullptr; base.mb = nullptr; }
    return base;
}

template <typename T>
int adjoint(tpcf_ctx *ctx, int kind, int64_t n_a, const T *xa, const T *ya, const T *za, const T *vala, const T *wa,
            int64_t n_b, const T *xb, const T *yb, const T *zb, const T *s1b, const T *s2b, const T *wb,
            const T *mb, const tpcf_bins *bins, uint32_t flags, const double *gbar1, const double *gbar2,
            double *dval_a, double *dw_a, double *d1_b, double *d2_b, double *dw_b, double *dm_b)
{
    if (!ctx) return TPCF_ERR_ARG;
    ctx->err.clear();
    TPCF_TRY(adj_check<T>(ctx, kind, n_a, xa, ya, za, vala, wa, n_b, xb, yb, zb, s1b, s2b, wb, mb, bins, flags));
    if (!gbar1 || !gbar2) return fail(ctx, TPCF_ERR_ARG, "upstream gradients are NULL");
    TPCF_HIP(ctx, hipSetDevice(ctx->device));
    const double t0 = now_ms();
    hipStream_t s = ctx->stream;
    const bool shear = (kind == SHEAR);
    const bool dev = (flags & TPCF_DEVICE_INPUTS) != 0;
    const uint32_t run_flags = flags | TPCF_DEVICE_INPUTS;

    CallArgs<T> base = adj_base<T>(kind, n_a, xa, ya, za, vala, wa, n_b, xb, yb, zb, s1b, s2b, wb, mb, bins, flags);
    if (!dev) {
        TPCF_TRY(upload_in(ctx, tpcf_ctx::B_IN0, xa, n_a, &base.xa));
        TPCF_TRY(upload_in(ctx, tpcf_ctx::B_IN1, ya, n_a, &base.ya));
        TPCF_TRY(upload_in(ctx, tpcf_ctx::B_IN2, shear ? za : (const T *)nullptr, n_a, &base.za));
        TPCF_TRY(upload_in(ctx, tpcf_ctx::B_IN3, vala, n_a, &base.vala));
        TPCF_TRY(upload_in(ctx, tpcf_ctx::B_IN4, wa, n_a, &base.wa));
        TPCF_TRY(upload_in(ctx, tpcf_ctx::B_IN5, xb, n_b, &base.xb));
        TPCF_TRY(upload_in(ctx, tpcf_ctx::B_IN6, yb, n_b, &base.yb));
        TPCF_TRY(upload_in(ctx, tpcf_ctx::B_IN7, shear ? zb : (const T *)nullptr, n_b, &base.zb));
        TPCF_TRY(upload_in(ctx, tpcf_ctx::B_IN8, s1b, n_b, &base.s1b));
        TPCF_TRY(upload_in(ctx, tpcf_ctx::B_IN9, s2b, n_b, &base.s2b));
        TPCF_TRY(upload_in(ctx, tpcf_ctx::B_IN10, wb, n_b, &base.wb));
        TPCF_TRY(upload_in(ctx, tpcf_ctx::B_IN11, shear ? mb : (const T *)nullptr, n_b, &base.mb));
    }

    // ---- gradient buffers (device), zero for masked objects ----
    double *gA, *gB;
    TPCF_TRY(ensure(ctx, tpcf_ctx::B_GRAD_A, sizeof(double) * 2 * std::max<int64_t>(n_a, 1), (void **)&gA));
    TPCF_TRY(ensure(ctx, tpcf_ctx::B_GRAD_B, sizeof(double) * 4 * std::max<int64_t>(n_b, 1), (void **)&gB));
    double *dvA = dev ? dval_a : gA, *dwA = dev ? dw_a : gA + n_a;
    double *d1B = dev ? d1_b : gB, *d2B = dev ? d2_b : gB + n_b, *dwB = dev ? dw_b : gB + 2 * n_b,
           *dmB = dev ? dm_b : gB + 3 * n_b;
    // host outputs: every device slot is computed and only the requested ones
    // copied; device outputs: only the requested ones are computed
    if (!shear) dmB = nullptr;
    double *zero_list[6] = {dvA, dwA, d1B, d2B, dwB, dmB};
    const int64_t zero_n[6] = {n_a, n_a, n_b, n_b, n_b, n_b};
    for (int q = 0; q < 6; ++q)
        if (zero_list[q] && zero_n[q] > 0) fill_f64(s, zero_list[q], zero_n[q]);

    LensSide ls;
    Coef c;
    if (flags & TPCF_SPHERICAL) {
        SphGeo g;
        TPCF_TRY(sph_geo<T>(ctx, base, g));
        TPCF_TRY((lens_pass<T, double>(ctx, base, g.d, run_flags, t0, ls)));
        c = coef_of(ls.fwd, bins->n_bin, gbar1, gbar2);
        TPCF_TRY(lens_grad<T>(ctx, ls, base, c, dvA, dwA));
        TPCF_TRY((src_pass<T, double>(ctx, base, g.d, g.rev_frames, run_flags, t0, c, d1B, d2B, dwB, dmB)));
    } else {
        TPCF_TRY((lens_pass<T, T>(ctx, base, base, run_flags, t0, ls)));
        c = coef_of(ls.fwd, bins->n_bin, gbar1, gbar2);
        TPCF_TRY(lens_grad<T>(ctx, ls, base, c, dvA, dwA));
        TPCF_TRY((src_pass<T, T>(ctx, base, base, nullptr, run_flags, t0, c, d1B, d2B, dwB, dmB)));
    }
    if (!dev) {
        double *dst[6] = {dval_a, dw_a, d1_b, d2_b, dw_b, shear ? dm_b : nullptr};
        for (int q = 0; q < 6; ++q)
            if (dst[q] && zero_n[q] > 0)
                TPCF_HIP(ctx, hipMemcpyAsync(dst[q], zero_list[q], sizeof(double) * zero_n[q], hipMemcpyDeviceToHost, s));
    }
    TPCF_HIP(ctx, hipStreamSynchronize(s));
    ctx->stats.ms_total = now_ms() - t0;
    return TPCF_OK;
}

}  // namespace
}  // namespace tpcf

// A recorded lens-side pass: the input pointers (caller-owned device arrays)
// and copies of the lens-side marks, the lens order and the forward sums, so
// that the gradient for any upstream gradient needs only the source-side pass.
struct tpcf_tape {
    int kind = 0;
    int real_bytes = 4;
    int device = 0;
    bool spherical = false;
    tpcf_bins bins{};
    int64_t n_a = 0, n_b = 0;
    const void *in[12] = {};   // x_A, y_A, z_A, val_A, w_A, x_B, y_B, z_B, s1_B, s2_B, w_B, m_B
    double *mk = nullptr;      // marks planes [n_bin][4][n_lens]
    uint32_t *lidx = nullptr;  // [n_lens]
    int64_t n_lens = 0;
    std::vector<double> fwd;   // [n_bin][3]
};

namespace tpcf {
namespace {

void tape_free(tpcf_tape *t)
{
    if (!t) return;
    (void)hipSetDevice(t->device);
    if (t->mk) (void)hipFree(t->mk);
    if (t->lidx) (void)hipFree(t->lidx);
    delete t;
}

template <typename T>
CallArgs<T> tape_base(const tpcf_tape &t)
{
    const T *const *in = reinterpret_cast<const T *const *>(t.in);
    return adj_base<T>(t.kind, t.n_a, in[0], in[1], in[2], in[3], in[4], t.n_b, in[5], in[6], in[7], in[8], in[9],
                       in[10], in[11], &t.bins, TPCF_DEVICE_INPUTS);
}

template <typename T>
int record(tpcf_ctx *ctx, int kind, int64_t n_a, const T *xa, const T *ya, const T *za, const T *vala, const T *wa,
           int64_t n_b, const T *xb, const T *yb, const T *zb, const T *s1b, const T *s2b, const T *wb, const T *mb,
           const tpcf_bins *bins, uint32_t flags, double *g1, double *g2, tpcf_tape **tape)
{
    if (!ctx) return TPCF_ERR_ARG;
    ctx->err.clear();
    if (!tape) return fail(ctx, TPCF_ERR_ARG, "tape is NULL");
    *tape = nullptr;
    TPCF_TRY(adj_check<T>(ctx, kind, n_a, xa, ya, za, vala, wa, n_b, xb, yb, zb, s1b, s2b, wb, mb, bins, flags));
    if (!(flags & TPCF_DEVICE_INPUTS))
        return fail(ctx, TPCF_ERR_ARG, "a recorded adjoint keeps the input pointers: TPCF_DEVICE_INPUTS is required");
    TPCF_HIP(ctx, hipSetDevice(ctx->device));
    const double t0 = now_ms();
    hipStream_t s = ctx->stream;
    tpcf_tape *t = new tpcf_tape;
    t->kind = kind;
    t->real_bytes = (int)sizeof(T);
    t->device = ctx->device;
    t->bins = *bins;
    t->n_a = n_a;
    t->n_b = n_b;
    t->spherical = (flags & TPCF_SPHERICAL) != 0;
    const T *in[12] = {xa, ya, za, vala, wa, xb, yb, zb, s1b, s2b, wb, mb};
    for (int q = 0; q < 12; ++q) t->in[q] = in[q];
    const CallArgs<T> base = tape_base<T>(*t);
    LensSide ls;
    int rc;
    if (t->spherical) {
        SphGeo g;
        rc = sph_geo<T>(ctx, base, g);
        if (rc == TPCF_OK) rc = lens_pass<T, double>(ctx, base, g.d, flags, t0, ls);
    } else {
        rc = lens_pass<T, T>(ctx, base, base, flags, t0, ls);
    }
    if (rc == TPCF_OK && ls.n_lens > 0) {
        const size_t nm = sizeof(double) * 4 * (size_t)ls.n_lens * bins->n_bin;
        if (hipMalloc((void **)&t->mk, nm) != hipSuccess || hipMalloc((void **)&t->lidx, 4 * ls.n_lens) != hipSuccess) {
            (void)hipGetLastError();
            rc = fail(ctx, TPCF_ERR_ALLOC, "tape: device allocation of the lens-side marks failed");
        } else {
            t->n_lens = ls.n_lens;
            const hipError_t e1 = hipMemcpyAsync(t->mk, ls.mk, nm, hipMemcpyDeviceToDevice, s);
            const hipError_t e2 = hipMemcpyAsync(t->lidx, ls.lidx, 4 * ls.n_lens, hipMemcpyDeviceToDevice, s);
            const hipError_t e3 = hipStreamSynchronize(s);
            if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess)
                rc = fail(ctx, TPCF_ERR_HIP, "tape: copy of the lens-side marks failed");
        }
    }
    if (rc != TPCF_OK) {
        tape_free(t);
        return rc;
    }
    t->fwd = ls.fwd;
    for (int b = 0; b < bins->n_bin; ++b) {
        const double G = t->fwd[3 * b], X = t->fwd[3 * b + 1], K = t->fwd[3 * b + 2];
        if (g1) g1[b] = K > 0.0 ? G / K : 0.0;
        if (g2) g2[b] = K > 0.0 ? X / K : 0.0;
    }
    ctx->stats.ms_total = now_ms() - t0;
    *tape = t;
    return TPCF_OK;
}

template <typename T>
int replay(tpcf_ctx *ctx, const tpcf_tape *t, const double *gbar1, const double *gbar2, double *dval_a,
           double *dw_a, double *d1_b, double *d2_b, double *dw_b, double *dm_b)
{
    TPCF_HIP(ctx, hipSetDevice(ctx->device));
    const double t0 = now_ms();
    hipStream_t s = ctx->stream;
    const CallArgs<T> base = tape_base<T>(*t);
    if (t->kind != SHEAR) dm_b = nullptr;
    double *outs[6] = {dval_a, dw_a, d1_b, d2_b, dw_b, dm_b};
    const int64_t n[6] = {t->n_a, t->n_a, t->n_b, t->n_b, t->n_b, t->n_b};
    for (int q = 0; q < 6; ++q)
        if (outs[q] && n[q] > 0) fill_f64(s, outs[q], n[q]);
    LensSide ls;
    ls.mk = t->mk;
    ls.lidx = t->lidx;
    ls.n_lens = t->n_lens;
    ls.fwd = t->fwd;
    const Coef c = coef_of(t->fwd, t->bins.n_bin, gbar1, gbar2);
    TPCF_TRY(lens_grad<T>(ctx, ls, base, c, dval_a, dw_a));
    if (t->spherical) {
        SphGeo g;
        TPCF_TRY(sph_geo<T>(ctx, base, g));
        TPCF_TRY((src_pass<T, double>(ctx, base, g.d, g.rev_frames, TPCF_DEVICE_INPUTS | TPCF_SPHERICAL, t0, c, d1_b,
                                      d2_b, dw_b, dm_b)));
    } else {
        TPCF_TRY((src_pass<T, T>(ctx, base, base, nullptr, TPCF_DEVICE_INPUTS, t0, c, d1_b, d2_b, dw_b, dm_b)));
    }
    TPCF_HIP(ctx, hipStreamSynchronize(s));
    ctx->stats.ms_total = now_ms() - t0;
    return TPCF_OK;
}

}  // namespace
}  // namespace tpcf

using namespace tpcf;

extern "C" {

int tpcf_adjoint_shear_f32(tpcf_ctx *ctx, int64_t n_A, const float *x_A, const float *y_A, const float *z_A,
                           const float *val_A, const float *w_A, int64_t n_B, const float *x_B, const float *y_B,
                           const float *z_B, const float *e1_B, const float *e2_B, const float *w_B,
                           const float *m_B, const tpcf_bins *bins, uint32_t flags, const double *gbar_t,
                           const double *gbar_x, double *val_A_bar, double *w_A_bar, double *e1_B_bar,
                           double *e2_B_bar, double *w_B_bar, double *m_B_bar)
{
    return adjoint<float>(ctx, SHEAR, n_A, x_A, y_A, z_A, val_A, w_A, n_B, x_B, y_B, z_B, e1_B, e2_B, w_B, m_B, bins,
                          flags, gbar_t, gbar_x, val_A_bar, w_A_bar, e1_B_bar, e2_B_bar, w_B_bar, m_B_bar);
}

int tpcf_adjoint_shear_f64(tpcf_ctx *ctx, int64_t n_A, const double *x_A, const double *y_A, const double *z_A,
                           const double *val_A, const double *w_A, int64_t n_B, const double *x_B, const double *y_B,
                           const double *z_B, const double *e1_B, const double *e2_B, const double *w_B,
                           const double *m_B, const tpcf_bins *bins, uint32_t flags, const double *gbar_t,
                           const double *gbar_x, double *val_A_bar, double *w_A_bar, double *e1_B_bar,
                           double *e2_B_bar, double *w_B_bar, double *m_B_bar)
{
    return adjoint<double>(ctx, SHEAR, n_A, x_A, y_A, z_A, val_A, w_A, n_B, x_B, y_B, z_B, e1_B, e2_B, w_B, m_B, bins,
                           flags, gbar_t, gbar_x, val_A_bar, w_A_bar, e1_B_bar, e2_B_bar, w_B_bar, m_B_bar);
}

int tpcf_adjoint_kappa_f32(tpcf_ctx *ctx, int64_t n_A, const float *x_A, const float *y_A, const float *val_A,
                           const float *w_A, int64_t n_B, const float *x_B, const float *y_B, const float *kappa_E,
                           const float *kappa_B, const float *w_B, const tpcf_bins *bins, uint32_t flags,
                           const double *gbar_E, const double *gbar_B, double *val_A_bar, double *w_A_bar,
                           double *kE_B_bar, double *kB_B_bar, double *w_B_bar)
{
    return adjoint<float>(ctx, KAPPA, n_A, x_A, y_A, nullptr, val_A, w_A, n_B, x_B, y_B, nullptr, kappa_E, kappa_B,
                          w_B, nullptr, bins, flags, gbar_E, gbar_B, val_A_bar, w_A_bar, kE_B_bar, kB_B_bar, w_B_bar,
                          nullptr);
}

int tpcf_adjoint_kappa_f64(tpcf_ctx *ctx, int64_t n_A, const double *x_A, const double *y_A, const double *val_A,
                           const double *w_A, int64_t n_B, const double *x_B, const double *y_B,
                           const double *kappa_E, const double *kappa_B, const double *w_B, const tpcf_bins *bins,
                           uint32_t flags, const double *gbar_E, const double *gbar_B, double *val_A_bar,
                           double *w_A_bar, double *kE_B_bar, double *kB_B_bar, double *w_B_bar)
{
    return adjoint<double>(ctx, KAPPA, n_A, x_A, y_A, nullptr, val_A, w_A, n_B, x_B, y_B, nullptr, kappa_E, kappa_B,
                           w_B, nullptr, bins, flags, gbar_E, gbar_B, val_A_bar, w_A_bar, kE_B_bar, kB_B_bar, w_B_bar,
                           nullptr);
}

int tpcf_adjoint_record_shear_f32(tpcf_ctx *ctx, int64_t n_A, const float *x_A, const float *y_A, const float *z_A,
                                  const float *val_A, const float *w_A, int64_t n_B, const float *x_B,
                                  const float *y_B, const float *z_B, const float *e1_B, const float *e2_B,
                                  const float *w_B, const float *m_B, const tpcf_bins *bins, uint32_t flags,
                                  double *g_t, double *g_x, tpcf_tape **tape)
{
    return record<float>(ctx, SHEAR, n_A, x_A, y_A, z_A, val_A, w_A, n_B, x_B, y_B, z_B, e1_B, e2_B, w_B, m_B, bins,
                         flags, g_t, g_x, tape);
}

int tpcf_adjoint_record_shear_f64(tpcf_ctx *ctx, int64_t n_A, const double *x_A, const double *y_A,
                                  const double *z_A, const double *val_A, const double *w_A, int64_t n_B,
                                  const double *x_B, const double *y_B, const double *z_B, const double *e1_B,
                                  const double *e2_B, const double *w_B, const double *m_B, const tpcf_bins *bins,
                                  uint32_t flags, double *g_t, double *g_x, tpcf_tape **tape)
{
    return record<double>(ctx, SHEAR, n_A, x_A, y_A, z_A, val_A, w_A, n_B, x_B, y_B, z_B, e1_B, e2_B, w_B, m_B, bins,
                          flags, g_t, g_x, tape);
}

int tpcf_adjoint_record_kappa_f32(tpcf_ctx *ctx, int64_t n_A, const float *x_A, const float *y_A, const float *val_A,
                                  const float *w_A, int64_t n_B, const float *x_B, const float *y_B,
                                  const float *kappa_E, const float *kappa_B, const float *w_B, const tpcf_bins *bins,
                                  uint32_t flags, double *xi_E, double *xi_B, tpcf_tape **tape)
{
    return record<float>(ctx, KAPPA, n_A, x_A, y_A, nullptr, val_A, w_A, n_B, x_B, y_B, nullptr, kappa_E, kappa_B,
                         w_B, nullptr, bins, flags, xi_E, xi_B, tape);
}

int tpcf_adjoint_record_kappa_f64(tpcf_ctx *ctx, int64_t n_A, const double *x_A, const double *y_A,
                                  const double *val_A, const double *w_A, int64_t n_B, const double *x_B,
                                  const double *y_B, const double *kappa_E, const double *kappa_B, const double *w_B,
                                  const tpcf_bins *bins, uint32_t flags, double *xi_E, double *xi_B, tpcf_tape **tape)
{
    return record<double>(ctx, KAPPA, n_A, x_A, y_A, nullptr, val_A, w_A, n_B, x_B, y_B, nullptr, kappa_E, kappa_B,
                          w_B, nullptr, bins, flags, xi_E, xi_B, tape);
}

int tpcf_adjoint_replay(tpcf_ctx *ctx, const tpcf_tape *tape, const double *gbar_1, const double *gbar_2,
                        double *val_A_bar, double *w_A_bar, double *s1_B_bar, double *s2_B_bar, double *w_B_bar,
                        double *m_B_bar)
{
    if (!ctx) return TPCF_ERR_ARG;
    ctx->err.clear();
    if (!tape) return fail(ctx, TPCF_ERR_ARG, "tape is NULL");
    if (!gbar_1 || !gbar_2) return fail(ctx, TPCF_ERR_ARG, "upstream gradients are NULL");
    if (tape->device != ctx->device) return fail(ctx, TPCF_ERR_ARG, "the tape was recorded on another device");
    if (tape->real_bytes == 4)
        return replay<float>(ctx, tape, gbar_1, gbar_2, val_A_bar, w_A_bar, s1_B_bar, s2_B_bar, w_B_bar, m_B_bar);
    return replay<double>(ctx, tape, gbar_1, gbar_2, val_A_bar, w_A_bar, s1_B_bar, s2_B_bar, w_B_bar, m_B_bar);
}

int tpcf_tape_destroy(tpcf_tape *tape)
{
    tape_free(tape);
    return TPCF_OK;
}

}  // extern "C"
