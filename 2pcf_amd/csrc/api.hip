// api.hip -- the C-ABI of tpcf_gpu.h: binning, context, and the four
// correlation entry points that stand in for shear_2pcf / kappa_2pcf
// (src/tpcf.f90:134-188) and their workers tangential_shear_tpcf /
// scalar_kappa_tpcf (src/tpcf.f90:190-545).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>

#include "internal.hpp"

using namespace tpcf;

namespace tpcf {

void fill_u64(hipStream_t s, unsigned long long *p, int64_t n);
void fill_f64(hipStream_t s, double *p, int64_t n);

int fail(tpcf_ctx *ctx, int code, const std::string &msg)
{
    if (ctx) ctx->err = msg;
    return code;
}

int ensure(tpcf_ctx *ctx, int which, size_t bytes, void **ptr)
{
    DevBuf &b = ctx->buf[which];
    bytes = std::max<size_t>(bytes, 256);
    if (ctx->arena) {
        // allocation-free call: a fresh region of the scratch array (the
        // buffer's old contents are not kept, as after a regrowth below)
        if (b.cap < bytes) {
            const size_t off = (ctx->arena_off + 255) & ~(size_t)255;
            if (off > ctx->arena_cap || bytes > ctx->arena_cap - off)
                return fail(ctx, TPCF_ERR_ALLOC,
                            "the call needs more device memory than the " + std::to_string(ctx->arena_cap >> 20) +
                                " MiB scratch region of the allocation-free entry point");
            b.p = ctx->arena + off;
            b.cap = bytes;
            ctx->arena_off = off + bytes;
        }
        *ptr = b.p;
        return TPCF_OK;
    }
    if (b.cap < bytes) {
        if (b.p) {
            hipError_t e = hipFree(b.p);
            b.p = nullptr;
            b.cap = 0;
            if (e != hipSuccess) return fail(ctx, TPCF_ERR_HIP, std::string("hipFree: ") + hipGetErrorString(e));
        }
        const size_t want = bytes + bytes / 4;
        if (hipMalloc(&b.p, want) != hipSuccess) {
            (void)hipGetLastError();
            b.p = nullptr;
            if (hipMalloc(&b.p, bytes) != hipSuccess) {
                (void)hipGetLastError();
                b.p = nullptr;
                return fail(ctx, TPCF_ERR_ALLOC, "device allocation of " + std::to_string(bytes) + " bytes failed");
            }
            b.cap = bytes;
        } else {
            b.cap = want;
        }
    }
    *ptr = b.p;
    return TPCF_OK;
}

double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

namespace {

// Sum of the per-XCD accumulator copies into the output layout
// [q][P_B][P_A][n_bin] (q = Theta, G_t, G_x, W, K) and [P_B][P_A][n_bin].
__global__ void k_reduce_pp(int32_t n_pa, int32_t n_pb, int32_t n_bin, const double *__restrict__ pp,
                            const unsigned long long *__restrict__ ppn, double *__restrict__ out,
                            long long *__restrict__ outn)
{
    const int64_t total = (int64_t)n_pa * n_pb * n_bin;
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int32_t bin = (int32_t)(t % n_bin);
    const int64_t r = t / n_bin;
    const int32_t pb = (int32_t)(r % n_pb), pa = (int32_t)(r / n_pb);
    double acc[kNQ] = {0, 0, 0, 0, 0};
    unsigned long long n = 0;
    for (int s = 0; s < kSlots; ++s) {
        const size_t o = (((size_t)s * n_pa + pa) * n_pb + pb) * n_bin + bin;
        for (int q = 0; q < kNQ; ++q) acc[q] += pp[o * kNQ + q];
        n += ppn[o];
    }
    const size_t d = ((size_t)pb * n_pa + pa) * n_bin + bin;
    for (int q = 0; q < kNQ; ++q) out[(size_t)q * total + d] = acc[q];
    outn[d] = (long long)n;
}

// Per-bin totals over all patch pairs: one block per bin.  tot holds the
// FP64 sums [n_bin][kNQ] followed by the pair counts [n_bin] (int64, summed
// as integers).
__global__ void k_bin_totals(int64_t n_pp, int32_t n_bin, const double *__restrict__ ppo,
                             const long long *__restrict__ ppno, double *__restrict__ tot)
{
    const int bin = blockIdx.x;
    const int64_t total = n_pp * n_bin;
    __shared__ double sh[kNQ][256];
    __shared__ unsigned long long shn[256];
    double acc[kNQ] = {0, 0, 0, 0, 0};
    unsigned long long n = 0;
    for (int64_t r = threadIdx.x; r < n_pp; r += blockDim.x) {
        const size_t d = (size_t)r * n_bin + bin;
        for (int q = 0; q < kNQ; ++q) acc[q] += ppo[(size_t)q * total + d];
        n += (unsigned long long)ppno[d];
    }
    for (int q = 0; q < kNQ; ++q) sh[q][threadIdx.x] = acc[q];
    shn[threadIdx.x] = n;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            for (int q = 0; q < kNQ; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + s];
            shn[threadIdx.x] += shn[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x < kNQ) tot[bin * kNQ + threadIdx.x] = sh[threadIdx.x][0];
    if (threadIdx.x == 0) reinterpret_cast<long long *>(tot + (size_t)n_bin * kNQ)[bin] = (long long)shn[0];
}

// Per-lens marks from sorted order back to the caller's order (columns of
// real(n_bin, n_A), src/tpcf.f90:204-205) plus marks_weight = sum of w_marks.
// Any of the five outputs may be NULL.
template <typename R>
struct MarksOut {
    R *q[5];
};

// The marks planes [n_bin][4][n_lens] (sorted lens order) to the caller's
// rows [n_A][n_bin] per component, through an LDS tile of 64 lenses x up to
// kMarksBins bins at a time: coalesced reads along the lenses, each lens' row
// written by consecutive threads (the rows themselves land in the caller's
// order, i.e. scattered).  marks_weight is the row sum of the W plane.
constexpr int kMarksTile = 64;
constexpr int kMarksBins = 32;

template <typename R>
__global__ __launch_bounds__(256) void k_marks_out(int64_t n_lens, int32_t n_bin, const uint32_t *__restrict__ lidx,
                                                   const double *__restrict__ mk, MarksOut<R> out)
{
    __shared__ double tile[kMarksBins][kMarksTile + 1];
    const int64_t i0 = (int64_t)blockIdx.x * kMarksTile;
    const int nl = (int)min<int64_t>(kMarksTile, n_lens - i0);
    double wsum = 0.0;   // marks_weight of lens i0 + threadIdx.x
    for (int q = 0; q < 4; ++q) {
        const bool want_w = (q == 3 && out.q[4] != nullptr);
        if (!out.q[q] && !want_w) continue;   // the same for the whole block
        for (int b0 = 0; b0 < n_bin; b0 += kMarksBins) {
            const int nb = min(kMarksBins, n_bin - b0);
            for (int e = threadIdx.x; e < nb * kMarksTile; e += blockDim.x) {
                const int b = e / kMarksTile, i = e % kMarksTile;
                tile[b][i] = (i < nl) ? mk[mark_at(n_lens, b0 + b, q, i0 + i)] : 0.0;
            }
            __syncthreads();
            if (out.q[q])
                for (int e = threadIdx.x; e < nl * nb; e += blockDim.x) {
                    const int i = e / nb, b = e - i * nb;
                    out.q[q][(size_t)lidx[i0 + i] * n_bin + b0 + b] = (R)tile[b][i];
                }
            if (want_w && (int)threadIdx.x < nl)
                for (int b = 0; b < nb; ++b) wsum += tile[b][threadIdx.x];
            __syncthreads();
        }
    }
    if (out.q[4] && (int)threadIdx.x < nl) out.q[4][lidx[i0 + threadIdx.x]] = (R)wsum;
}

// The K plane of every bin (blockIdx.y) = its W plane.
__global__ void k_marks_k_from_w(int64_t n_lens, double *__restrict__ mk)
{
    const int b = blockIdx.y;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n_lens; i += (int64_t)gridDim.x * blockDim.x)
        mk[mark_at(n_lens, b, 2, i)] = mk[mark_at(n_lens, b, 3, i)];
}

template <typename R>
__global__ void k_zero(R *p, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = (R)0;
}

// Zero the marks outputs (the caller's lenses outside the pair set keep 0)
// and write the pair set's marks into them.
template <typename R>
int marks_to(tpcf_ctx *ctx, hipStream_t s, int64_t n_lens, int32_t n_bin, int64_t n_a, const uint32_t *lidx,
             const double *mk, const MarksOut<R> &out)
{
    for (int q = 0; q < 5; ++q) {
        const int64_t n = (q < 4) ? n_a * n_bin : n_a;
        if (!out.q[q] || n == 0) continue;
        hipLaunchKernelGGL(k_zero<R>, dim3((unsigned)std::min<int64_t>(8192, (n + 255) / 256)), dim3(256), 0, s,
                           out.q[q], n);
    }
    if (n_lens > 0)
        hipLaunchKernelGGL(k_marks_out<R>, dim3((unsigned)((n_lens + kMarksTile - 1) / kMarksTile)), dim3(256), 0, s,
                           n_lens, n_bin, lidx, mk, out);
    TPCF_HIP(ctx, hipGetLastError());
    return TPCF_OK;
}

void put_real(void *dst, int rb, int i, double v)
{
    if (!dst) return;
    if (rb == 8) static_cast<double *>(dst)[i] = v;
    else static_cast<float *>(dst)[i] = (float)v;
}

}  // namespace

int launch_marks_out(tpcf_ctx *ctx, hipStream_t s, int64_t n_lens, int32_t n_bin, int64_t n_a, const uint32_t *lidx,
                     double *mk, void *const out[5], int rb, bool k_from_w)
{
    if (k_from_w && n_lens > 0) {
        hipLaunchKernelGGL(k_marks_k_from_w, dim3((unsigned)std::min<int64_t>(4096, (n_lens + 255) / 256), (unsigned)n_bin),
                           dim3(256), 0, s, n_lens, mk);
        TPCF_HIP(ctx, hipGetLastError());
    }
    if (rb == 8)
        return marks_to<double>(ctx, s, n_lens, n_bin, n_a, lidx, mk,
                                MarksOut<double>{{(double *)out[0], (double *)out[1], (double *)out[2],
                                                  (double *)out[3], (double *)out[4]}});
    return marks_to<float>(ctx, s, n_lens, n_bin, n_a, lidx, mk,
                           MarksOut<float>{{(float *)out[0], (float *)out[1], (float *)out[2], (float *)out[3],
                                            (float *)out[4]}});
}

int check_bins(tpcf_ctx *ctx, const tpcf_bins *b)
{
    if (!b) return fail(ctx, TPCF_ERR_ARG, "bins is NULL");
    if (b->n_bin < 1 || b->n_bin > TPCF_MAX_BIN)
        return fail(ctx, TPCF_ERR_BINS, "n_bin must be in [1, " + std::to_string(TPCF_MAX_BIN) + "]");
    for (int i = 0; i <= b->n_bin; ++i)
        if (!std::isfinite(b->edges[i]) || b->edges[i] < 0.0 || (i > 0 && !(b->edges[i] > b->edges[i - 1])))
            return fail(ctx, TPCF_ERR_BINS, "bin edges must be finite, non-negative and increasing");
    return TPCF_OK;
}

namespace {

template <typename T>
int upload(tpcf_ctx *ctx, int which, const T *src, int64_t n, const T **dst)
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

}  // namespace

template <typename T>
int correlate(tpcf_ctx *ctx, int kind, int64_t n_a, const T *xa, const T *ya, const T *za, const T *vala,
              const T *wa, int64_t n_b, const T *xb, const T *yb, const T *zb, const T *s1b, const T *s2b,
              const T *wb, const T *mb, const tpcf_bins *bins, uint32_t flags, const int32_t *pat_a,
              const int32_t *pat_b, int32_t n_pa, int32_t n_pb, tpcf_out *out)
{
    if (!ctx) return TPCF_ERR_ARG;
    ctx->err.clear();
    TPCF_TRY(check_bins(ctx, bins));
    if (!out) return fail(ctx, TPCF_ERR_ARG, "out is NULL");
    if (out->real_bytes != 4 && out->real_bytes != 8) return fail(ctx, TPCF_ERR_ARG, "out->real_bytes must be 4 or 8");
    if (flags & ~(uint32_t)(TPCF_DEVICE_INPUTS | TPCF_VERBOSE | TPCF_NO_KD_TREE | TPCF_SPHERICAL | TPCF_DEVICE_MARKS))
        return fail(ctx, TPCF_ERR_ARG, "unknown flag");
    if (n_a < 0 || n_b < 0 || n_a >= ((int64_t)1 << 31) || n_b >= ((int64_t)1 << 31))
        return fail(ctx, TPCF_ERR_ARG, "catalogue sizes must be in [0, 2^31)");
    const bool shear = (kind == SHEAR);
    if (n_a > 0 && (!xa || !ya || !vala || !wa || (shear && !za)))
        return fail(ctx, TPCF_ERR_ARG, "a required lens array is NULL");
    if (n_b > 0 && (!xb || !yb || !s1b || !s2b || !wb || (shear && (!zb || !mb))))
        return fail(ctx, TPCF_ERR_ARG, "a required source array is NULL");
    if (!pat_a) n_pa = 1;
    if (!pat_b) n_pb = 1;
    if (n_pa < 1 || n_pb < 1) return fail(ctx, TPCF_ERR_ARG, "n_patch must be >= 1");
    if ((int64_t)n_pa * n_pb * bins->n_bin * kSlots > ((int64_t)1 << 31))
        return fail(ctx, TPCF_ERR_ARG, "too many patch pairs");
    TPCF_HIP(ctx, hipSetDevice(ctx->device));

    const double t0 = now_ms();
    hipStream_t s = ctx->stream;

    CallArgs<T> a;
    a.kind = kind;
    a.n_a = n_a;
    a.n_b = n_b;
    a.n_pa = n_pa;
    a.n_pb = n_pb;
    a.bins = bins;
    a.marks = (out->xi_E_marks || out->xi_B_marks || out->K_marks || out->w_marks || out->marks_weight);
    a.verbose = (flags & TPCF_VERBOSE) != 0;
    TPCF_HIP(ctx, hipEventRecord(ctx->ev[0], s));
    if (flags & TPCF_DEVICE_INPUTS) {
        a.xa = xa; a.ya = ya; a.za = za; a.vala = vala; a.wa = wa;
        a.xb = xb; a.yb = yb; a.zb = zb; a.s1b = s1b; a.s2b = s2b; a.wb = wb; a.mb = mb;
        a.patch_a = pat_a; a.patch_b = pat_b;
    } else {
        TPCF_TRY(upload(ctx, tpcf_ctx::B_IN0, xa, n_a, &a.xa));
        TPCF_TRY(upload(ctx, tpcf_ctx::B_IN1, ya, n_a, &a.ya));
        TPCF_TRY(upload(ctx, tpcf_ctx::B_IN2, shear ? za : (const T *)nullptr, n_a, &a.za));
        TPCF_TRY(upload(ctx, tpcf_ctx::B_IN3, vala, n_a, &a.vala));
        TPCF_TRY(upload(ctx, tpcf_ctx::B_IN4, wa, n_a, &a.wa));
        TPCF_TRY(upload(ctx, tpcf_ctx::B_IN5, xb, n_b, &a.xb));
        TPCF_TRY(upload(ctx, tpcf_ctx::B_IN6, yb, n_b, &a.yb));
        TPCF_TRY(upload(ctx, tpcf_ctx::B_IN7, shear ? zb : (const T *)nullptr, n_b, &a.zb));
        TPCF_TRY(upload(ctx, tpcf_ctx::B_IN8, s1b, n_b, &a.s1b));
        TPCF_TRY(upload(ctx, tpcf_ctx::B_IN9, s2b, n_b, &a.s2b));
        TPCF_TRY(upload(ctx, tpcf_ctx::B_IN10, wb, n_b, &a.wb));
        TPCF_TRY(upload(ctx, tpcf_ctx::B_IN11, shear ? mb : (const T *)nullptr, n_b, &a.mb));
        TPCF_TRY(upload(ctx, tpcf_ctx::B_PATCH_A, pat_a, n_a, &a.patch_a));
        TPCF_TRY(upload(ctx, tpcf_ctx::B_PATCH_B, pat_b, n_b, &a.patch_b));
    }
    if (!shear) { a.za = nullptr; a.zb = nullptr; a.mb = nullptr; }
    if (flags & TPCF_SPHERICAL) return sph_correlate<T>(ctx, a, flags, out, t0);
    return run_call<T>(ctx, a, flags, out, t0, nullptr);
}

template int correlate<float>(tpcf_ctx *, int, int64_t, const float *, const float *, const float *, const float *,
                              const float *, int64_t, const float *, const float *, const float *, const float *,
                              const float *, const float *, const float *, const tpcf_bins *, uint32_t,
                              const int32_t *, const int32_t *, int32_t, int32_t, tpcf_out *);
template int correlate<double>(tpcf_ctx *, int, int64_t, const double *, const double *, const double *,
                               const double *, const double *, int64_t, const double *, const double *,
                               const double *, const double *, const double *, const double *, const double *,
                               const tpcf_bins *, uint32_t, const int32_t *, const int32_t *, int32_t, int32_t,
                               tpcf_out *);

// Everything after the inputs are on the device: preparation, the pair
// kernel, reductions and outputs.  With a.marks_device the marks stay in the
// context's FP64 buffer (sorted lens order) for the adjoint; *pout receives the
// preparation (lens order) when given.
template <typename T>
int run_call(tpcf_ctx *ctx, CallArgs<T> &a, uint32_t flags, tpcf_out *out, double t0, Prepared<T> *pout)
{
    const int64_t n_a = a.n_a, n_b = a.n_b;
    const int32_t n_pa = a.n_pa, n_pb = a.n_pb;
    const bool shear = (a.kind == SHEAR);
    const tpcf_bins *bins = a.bins;
    const int n_bin = bins->n_bin;
    const int rb = out->real_bytes;
    hipStream_t s = ctx->stream;
    tpcf_stats st{};
    TPCF_HIP(ctx, hipEventRecord(ctx->ev[1], s));   // inputs on the device (ev[0]: before the uploads)

    std::vector<double> tot((size_t)n_bin * (kNQ + 1), 0.0);
    const int64_t n_pp = (int64_t)n_pa * n_pb;
    const int64_t pp_total = n_pp * n_bin;
    const bool want_pp = out->pp_n || out->pp_theta || out->pp_gt || out->pp_gx || out->pp_w || out->pp_k;
    std::vector<double> hpp;
    std::vector<long long> hppn;
    const bool empty = (n_a == 0 || n_b == 0);
    Prepared<T> p;
    std::memset(&p, 0, sizeof p);
    if (!empty) {
        TPCF_TRY(prepare<T>(ctx, a, p));
        if constexpr (sizeof(T) == 8) {
            if (a.sph_pa) TPCF_TRY(sph_finish(ctx, a, p));
        }
        double *pp, *marks = nullptr;
        unsigned long long *ppn;
        TPCF_TRY(ensure(ctx, tpcf_ctx::B_PP, sizeof(double) * kNQ * kSlots * pp_total, (void **)&pp));
        TPCF_TRY(ensure(ctx, tpcf_ctx::B_PPN, sizeof(unsigned long long) * kSlots * pp_total, (void **)&ppn));
        fill_f64(s, pp, kNQ * kSlots * pp_total);
        fill_u64(s, ppn, kSlots * pp_total);
        if (a.marks) {
            TPCF_TRY(ensure(ctx, tpcf_ctx::B_MARKS, sizeof(double) * 4 * std::max<int64_t>(p.n_lens, 1) * n_bin,
                            (void **)&marks));
            fill_f64(s, marks, 4 * p.n_lens * n_bin);
        }
        TPCF_TRY(run_pairs<T>(ctx, a, p, pp, ppn, marks));
        if (a.marks && p.n_lens > 0 && (a.kind == KAPPA || p.m_zero)) {
            // the pair kernel leaves K marks to this copy when K == W
            hipLaunchKernelGGL(k_marks_k_from_w,
                               dim3((unsigned)std::min<int64_t>(4096, (p.n_lens + 255) / 256), (unsigned)n_bin),
                               dim3(256), 0, s, p.n_lens, marks);
            TPCF_HIP(ctx, hipGetLastError());
        }
        if (p.n_items == 0) {
            TPCF_HIP(ctx, hipEventRecord(ctx->ev[2], s));
            TPCF_HIP(ctx, hipEventRecord(ctx->ev[3], s));
        }
        // ---- reduce, totals ----
        double *ppo, *dtot;
        long long *ppno;
        void *vb;
        TPCF_TRY(ensure(ctx, tpcf_ctx::B_OUT, sizeof(double) * (kNQ + 1) * pp_total + sizeof(double) * n_bin * (kNQ + 1) + 64,
                        &vb));
        ppo = static_cast<double *>(vb);
        ppno = reinterpret_cast<long long *>(ppo + (size_t)kNQ * pp_total);
        dtot = reinterpret_cast<double *>(ppno + pp_total);
        TPCF_TRY(launch_reduce(ctx, s, n_pa, n_pb, n_bin, pp, ppn, ppo, ppno, dtot));
        TPCF_HIP(ctx, hipMemcpyAsync(tot.data(), dtot, sizeof(double) * tot.size(), hipMemcpyDeviceToHost, s));
        if (want_pp) {
            hpp.resize((size_t)kNQ * pp_total);
            hppn.resize(pp_total);
            TPCF_HIP(ctx, hipMemcpyAsync(hpp.data(), ppo, sizeof(double) * hpp.size(), hipMemcpyDeviceToHost, s));
            TPCF_HIP(ctx, hipMemcpyAsync(hppn.data(), ppno, sizeof(long long) * hppn.size(), hipMemcpyDeviceToHost, s));
        }
        if (a.marks && !a.marks_device) {
            void *dsts[5] = {out->xi_E_marks, out->xi_B_marks, out->K_marks, out->w_marks, out->marks_weight};
            const size_t plane = (size_t)n_a * n_bin;
            void *mo = nullptr;
            if (!(flags & TPCF_DEVICE_MARKS))
                TPCF_TRY(ensure(ctx, tpcf_ctx::B_TMP, (size_t)rb * (4 * plane + n_a), &mo));
            // with TPCF_DEVICE_MARKS straight into the caller's arrays, else
            // into one staging buffer copied out below
            void *q5[5];
            for (int q = 0; q < 5; ++q)
                q5[q] = !dsts[q] ? nullptr : mo ? static_cast<char *>(mo) + (size_t)rb * q * plane : dsts[q];
            if (rb == 8)
                TPCF_TRY(marks_to<double>(ctx, s, p.n_lens, n_bin, n_a, p.lidx, marks,
                                          MarksOut<double>{{(double *)q5[0], (double *)q5[1], (double *)q5[2],
                                                            (double *)q5[3], (double *)q5[4]}}));
            else
                TPCF_TRY(marks_to<float>(ctx, s, p.n_lens, n_bin, n_a, p.lidx, marks,
                                         MarksOut<float>{{(float *)q5[0], (float *)q5[1], (float *)q5[2],
                                                          (float *)q5[3], (float *)q5[4]}}));
            for (int q = 0; q < 5 && mo; ++q) {
                if (!dsts[q]) continue;
                const size_t cnt = (q < 4) ? plane : (size_t)n_a;
                if (cnt == 0) continue;
                TPCF_HIP(ctx, hipMemcpyAsync(dsts[q], q5[q], (size_t)rb * cnt, hipMemcpyDeviceToHost, s));
            }
        }
    } else {
        TPCF_HIP(ctx, hipEventRecord(ctx->ev[2], s));
        TPCF_HIP(ctx, hipEventRecord(ctx->ev[3], s));
        if (want_pp) {
            hpp.assign((size_t)kNQ * pp_total, 0.0);
            hppn.assign(pp_total, 0);
        }
        const size_t plane = (size_t)n_a * n_bin;
        void *dsts[5] = {out->xi_E_marks, out->xi_B_marks, out->K_marks, out->w_marks, out->marks_weight};
        if (a.marks && !a.marks_device && (flags & TPCF_DEVICE_MARKS)) {
            if (rb == 8)
                TPCF_TRY(marks_to<double>(ctx, s, 0, n_bin, n_a, nullptr, nullptr,
                                          MarksOut<double>{{(double *)dsts[0], (double *)dsts[1], (double *)dsts[2],
                                                            (double *)dsts[3], (double *)dsts[4]}}));
            else
                TPCF_TRY(marks_to<float>(ctx, s, 0, n_bin, n_a, nullptr, nullptr,
                                         MarksOut<float>{{(float *)dsts[0], (float *)dsts[1], (float *)dsts[2],
                                                          (float *)dsts[3], (float *)dsts[4]}}));
        } else {
            for (int q = 0; q < 5 && !a.marks_device; ++q)
                if (dsts[q]) std::memset(dsts[q], 0, (size_t)rb * ((q < 4) ? plane : (size_t)n_a));
        }
    }
    TPCF_HIP(ctx, hipEventRecord(ctx->ev[4], s));
    TPCF_HIP(ctx, hipStreamSynchronize(s));

    // ---- per-bin outputs and normalisation (src/tpcf.f90:369-376 / :536-543) ----
    int64_t npairs = 0;
    for (int b = 0; b < n_bin; ++b) {
        const double *t = &tot[(size_t)b * kNQ];
        const double th = t[0], gt = t[1], gx = t[2], w = t[3], k = t[4];
        int64_t n;
        std::memcpy(&n, &tot[(size_t)n_bin * kNQ + b], sizeof n);   // the int64 counts follow the sums
        npairs += n;
        put_real(out->theta, rb, b, (double)bins->bin_centers[b]);
        put_real(out->mean_theta, rb, b, th / w);   // NaN for an empty bin, as the reference
        put_real(out->g_t, rb, b, k > 0.0 ? gt / k : 0.0);
        put_real(out->g_x, rb, b, k > 0.0 ? gx / k : 0.0);
        put_real(out->weights, rb, b, w);
        if (out->n_pairs) out->n_pairs[b] = n;
        if (out->sum_theta) out->sum_theta[b] = th;
        if (out->sum_gt) out->sum_gt[b] = gt;
        if (out->sum_gx) out->sum_gx[b] = gx;
        if (out->sum_w) out->sum_w[b] = w;
        if (out->sum_k) out->sum_k[b] = k;
    }
    if (want_pp) {
        double *dst[kNQ] = {out->pp_theta, out->pp_gt, out->pp_gx, out->pp_w, out->pp_k};
        for (int q = 0; q < kNQ; ++q)
            if (dst[q]) std::memcpy(dst[q], hpp.data() + (size_t)q * pp_total, sizeof(double) * pp_total);
        if (out->pp_n) std::memcpy(out->pp_n, hppn.data(), sizeof(long long) * pp_total);
    }

    float ms = 0.f;
    if (!(flags & TPCF_DEVICE_INPUTS)) {
        TPCF_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
        st.ms_upload = ms;
    }
    TPCF_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev[1], ctx->ev[2]));
    st.ms_prepare = ms;
    TPCF_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]));
    st.ms_pairs = ms;
    TPCF_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev[3], ctx->ev[4]));
    st.ms_finish = ms;
    st.n_pairs = npairs;
    st.n_lens_used = p.n_lens;
    st.n_source_used = p.n_src;
    st.n_chunks = p.n_chunks;
    st.n_items = p.n_items;
    st.n_segments = p.n_segs;
    st.cell_size = p.grid.cell;
    if (ctx->opt_counters && !empty) {
        unsigned long long *probe = static_cast<unsigned long long *>(ctx->buf[tpcf_ctx::B_PROBE].p);
        unsigned long long h[8] = {};
        if (probe) TPCF_HIP(ctx, hipMemcpy(h, probe, sizeof h, hipMemcpyDeviceToHost));
        for (int c = 0; c < 3; ++c) {
            st.seg_visits[c] = (int64_t)h[c];
            st.pair_visits[c] = (int64_t)h[3 + c];
        }
    }
    st.ms_total = now_ms() - t0;
    ctx->stats = st;
    if (a.verbose) {
        std::printf("tpcf_gpu: %s, %lld lenses (%lld used) x %lld sources (%lld used), %d bins, %d x %d patches\n",
                    shear ? "tangential shear" : "kappa", (long long)n_a, (long long)p.n_lens, (long long)n_b,
                    (long long)p.n_src, n_bin, n_pa, n_pb);
        std::printf("tpcf_gpu: cell %.4g, %lld chunks, %lld segments, %lld pairs, pair kernel %.3f ms, total %.3f ms\n",
                    p.grid.cell, (long long)p.n_chunks, (long long)p.n_segs, (long long)npairs, st.ms_pairs, st.ms_total);
        std::fflush(stdout);
    }
    if (pout) *pout = p;
    return TPCF_OK;
}

int launch_reduce(tpcf_ctx *ctx, hipStream_t s, int32_t n_pa, int32_t n_pb, int32_t n_bin, const double *pp,
                  const unsigned long long *ppn, double *ppo, long long *ppno, double *tot)
{
    const int64_t n_pp = (int64_t)n_pa * n_pb, pp_total = n_pp * n_bin;
    hipLaunchKernelGGL(k_reduce_pp, dim3((unsigned)((pp_total + 255) / 256)), dim3(256), 0, s, n_pa, n_pb, n_bin, pp,
                       ppn, ppo, ppno);
    hipLaunchKernelGGL(k_bin_totals, dim3(n_bin), dim3(256), 0, s, n_pp, n_bin, ppo, ppno, tot);
    TPCF_HIP(ctx, hipGetLastError());
    return TPCF_OK;
}

template int run_call<float>(tpcf_ctx *, CallArgs<float> &, uint32_t, tpcf_out *, double, Prepared<float> *);
template int run_call<double>(tpcf_ctx *, CallArgs<double> &, uint32_t, tpcf_out *, double, Prepared<double> *);

}  // namespace tpcf

extern "C" {

int tpcf_create_bins(float x_min, float x_max, int32_t n_bin, int32_t spacing, float ihs_theta, tpcf_bins *bins)
{
    if (!bins) return TPCF_ERR_ARG;
    std::memset(bins, 0, sizeof *bins);
    if (n_bin < 1 || n_bin > TPCF_MAX_BIN) return TPCF_ERR_BINS;
    if (spacing < TPCF_LIN || spacing > TPCF_SQRT) return TPCF_ERR_BINS;   // src/utils.f90:152-159
    if (!std::isfinite(x_min) || !std::isfinite(x_max) || !(x_max > x_min) || !(x_min >= 0.0f)) return TPCF_ERR_BINS;
    if (spacing == TPCF_LOG && !(x_min > 0.0f)) return TPCF_ERR_BINS;
    const float p = (spacing == TPCF_IHS) ? ihs_theta : 1.0f;              // src/utils.f90:160-164
    if (spacing == TPCF_IHS && !(p > 0.0f)) return TPCF_ERR_BINS;
    bins->x_min = x_min;
    bins->x_max = x_max;
    bins->n_bin = n_bin;
    bins->spacing = spacing;
    bins->ihs_theta = p;
    // bin_size in the transformed space, as binning%bin_size (src/utils.f90:183-191)
    auto ihs = [](float x, float q) { return std::log(q * x + std::sqrt((q * x) * (q * x) + 1.0f)) / q; };
    switch (spacing) {
    case TPCF_LIN: bins->bin_size = (x_max - x_min) / (float)n_bin; break;
    case TPCF_LOG: bins->bin_size = (std::log10(x_max) - std::log10(x_min)) / (float)n_bin; break;
    case TPCF_IHS: bins->bin_size = (ihs(x_max, p) - ihs(x_min, p)) / (float)n_bin; break;
    default: bins->bin_size = (std::sqrt(x_max) - std::sqrt(x_min)) / (float)n_bin; break;
    }
    // FP64 edges: inverse transform of T(x_min) + i * (T(x_max) - T(x_min)) / n_bin;
    // the log spacing uses log2/exp2 with x_min factored out.
    const double lo = (double)x_min, hi = (double)x_max, pd = (double)p;
    bins->edges[0] = lo;
    bins->edges[n_bin] = hi;
    if (spacing == TPCF_LOG) {
        const double span = std::log2(hi / lo);
        for (int i = 1; i < n_bin; ++i) bins->edges[i] = lo * std::exp2(((double)i * span) / (double)n_bin);
    } else {
        auto fwd = [&](double x) {
            switch (spacing) {
            case TPCF_LIN: return x;
            case TPCF_IHS: return std::asinh(pd * x) / pd;
            default: return std::sqrt(x);
            }
        };
        auto inv = [&](double t) {
            switch (spacing) {
            case TPCF_LIN: return t;
            case TPCF_IHS: return std::sinh(pd * t) / pd;
            default: return t * t;
            }
        };
        const double t0 = fwd(lo), t1 = fwd(hi);
        for (int i = 1; i < n_bin; ++i) bins->edges[i] = inv(t0 + ((double)i * (t1 - t0)) / (double)n_bin);
    }
    for (int i = 1; i <= n_bin; ++i)
        if (!(bins->edges[i] > bins->edges[i - 1])) return TPCF_ERR_BINS;
    // arithmetic midpoints, also for log bins (src/utils.f90:229)
    for (int i = 0; i < n_bin; ++i)
        bins->bin_centers[i] = (float)(bins->edges[i] + (bins->edges[i + 1] - bins->edges[i]) / 2.0);
    return TPCF_OK;
}

int32_t tpcf_find_bin_index(double x, const tpcf_bins *bins)
{
    if (!bins || bins->n_bin < 1 || bins->n_bin > TPCF_MAX_BIN) return -1;
    const int n = bins->n_bin;
    const double d2 = x * x;
    const double lo = bins->edges[0] * bins->edges[0], hi = bins->edges[n] * bins->edges[n];
    if (!(d2 >= lo) || d2 > hi) return -1;
    if (d2 == hi) return n - 1;
    int k = 0;
    while (k + 1 < n && d2 >= bins->edges[k + 1] * bins->edges[k + 1]) ++k;
    return k;
}

int tpcf_ctx_create(int32_t device, tpcf_ctx **out)
{
    if (!out) return TPCF_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        return TPCF_ERR_NODEVICE;
    }
    if (device < 0) {
        if (hipGetDevice(&device) != hipSuccess) return TPCF_ERR_NODEVICE;
    }
    if (device >= n) return TPCF_ERR_NODEVICE;
    tpcf_ctx *ctx = new (std::nothrow) tpcf_ctx();
    if (!ctx) return TPCF_ERR_ALLOC;
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&ctx->own_stream) != hipSuccess) {
        delete ctx;
        return TPCF_ERR_HIP;
    }
    ctx->stream = ctx->own_stream;
    for (auto &e : ctx->ev)
        if (hipEventCreate(&e) != hipSuccess) {
            tpcf_ctx_destroy(ctx);
            return TPCF_ERR_HIP;
        }
    *out = ctx;
    return TPCF_OK;
}

int tpcf_ctx_destroy(tpcf_ctx *ctx)
{
    if (!ctx) return TPCF_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto &b : ctx->buf)
        if (b.p) (void)hipFree(b.p);
    for (auto &e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
    delete ctx;
    return TPCF_OK;
}

int tpcf_set_option(tpcf_ctx *ctx, int32_t option, double value)
{
    if (!ctx) return TPCF_ERR_ARG;
    switch (option) {
    case TPCF_OPT_CELL_SIZE:
        if (!(value >= 0.0) || !std::isfinite(value)) return fail(ctx, TPCF_ERR_ARG, "cell size must be >= 0");
        ctx->opt_cell = value;
        return TPCF_OK;
    case TPCF_OPT_COUNTERS:
        ctx->opt_counters = (value != 0.0);
        return TPCF_OK;
    case TPCF_OPT_SPLIT:
        if (!(value > 0.0) || !std::isfinite(value)) return fail(ctx, TPCF_ERR_ARG, "split must be > 0");
        ctx->opt_split = value;
        return TPCF_OK;
    case TPCF_OPT_CELL_TARGET:
        if (!(value > 0.0) || !std::isfinite(value)) return fail(ctx, TPCF_ERR_ARG, "cell target must be > 0");
        ctx->opt_cell_target = value;
        return TPCF_OK;
    default:
        return fail(ctx, TPCF_ERR_ARG, "unknown option");
    }
}

int tpcf_set_stream(tpcf_ctx *ctx, void *hip_stream)
{
    if (!ctx) return TPCF_ERR_ARG;
    ctx->stream = hip_stream ? static_cast<hipStream_t>(hip_stream) : ctx->own_stream;
    return TPCF_OK;
}

int tpcf_get_stats(const tpcf_ctx *ctx, tpcf_stats *stats)
{
    if (!ctx || !stats) return TPCF_ERR_ARG;
    *stats = ctx->stats;
    return TPCF_OK;
}

const char *tpcf_last_error(const tpcf_ctx *ctx) { return ctx ? ctx->err.c_str() : "no context"; }

int tpcf_shear_f32(tpcf_ctx *ctx, int64_t n_A, const float *x_A, const float *y_A, const float *z_A,
                   const float *val_A, const float *w_A, int64_t n_B, const float *x_B, const float *y_B,
                   const float *z_B, const float *e1_B, const float *e2_B, const float *w_B, const float *m_B,
                   const tpcf_bins *bins, uint32_t flags, const int32_t *patch_A, const int32_t *patch_B,
                   int32_t n_patch_A, int32_t n_patch_B, tpcf_out *out)
{
    return correlate<float>(ctx, SHEAR, n_A, x_A, y_A, z_A, val_A, w_A, n_B, x_B, y_B, z_B, e1_B, e2_B, w_B, m_B, bins,
                            flags, patch_A, patch_B, n_patch_A, n_patch_B, out);
}

int tpcf_kappa_f32(tpcf_ctx *ctx, int64_t n_A, const float *x_A, const float *y_A, const float *val_A,
                   const float *w_A, int64_t n_B, const float *x_B, const float *y_B, const float *kappa_E,
                   const float *kappa_B, const float *w_B, const tpcf_bins *bins, uint32_t flags,
                   const int32_t *patch_A, const int32_t *patch_B, int32_t n_patch_A, int32_t n_patch_B,
                   tpcf_out *out)
{
    return correlate<float>(ctx, KAPPA, n_A, x_A, y_A, nullptr, val_A, w_A, n_B, x_B, y_B, nullptr, kappa_E, kappa_B,
                            w_B, nullptr, bins, flags, patch_A, patch_B, n_patch_A, n_patch_B, out);
}

int tpcf_shear_f64(tpcf_ctx *ctx, int64_t n_A, const double *x_A, const double *y_A, const double *z_A,
                   const double *val_A, const double *w_A, int64_t n_B, const double *x_B, const double *y_B,
                   const double *z_B, const double *e1_B, const double *e2_B, const double *w_B, const double *m_B,
                   const tpcf_bins *bins, uint32_t flags, const int32_t *patch_A, const int32_t *patch_B,
                   int32_t n_patch_A, int32_t n_patch_B, tpcf_out *out)
{
    return correlate<double>(ctx, SHEAR, n_A, x_A, y_A, z_A, val_A, w_A, n_B, x_B, y_B, z_B, e1_B, e2_B, w_B, m_B, bins,
                             flags, patch_A, patch_B, n_patch_A, n_patch_B, out);
}

int tpcf_kappa_f64(tpcf_ctx *ctx, int64_t n_A, const double *x_A, const double *y_A, const double *val_A,
                   const double *w_A, int64_t n_B, const double *x_B, const double *y_B, const double *kappa_E,
                   const double *kappa_B, const double *w_B, const tpcf_bins *bins, uint32_t flags,
                   const int32_t *patch_A, const int32_t *patch_B, int32_t n_patch_A, int32_t n_patch_B,
                   tpcf_out *out)
{
    return correlate<double>(ctx, KAPPA, n_A, x_A, y_A, nullptr, val_A, w_A, n_B, x_B, y_B, nullptr, kappa_E, kappa_B,
                             w_B, nullptr, bins, flags, patch_A, patch_B, n_patch_A, n_patch_B, out);
}

}  // extern "C"
