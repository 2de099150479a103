// bootstrap.hip -- block bootstrap of the per-patch sums and the covariance of
// the resampled estimators: the resampling loop of
// survey_marked_point_block_bootstrap (src/tpcf.f90:1165-1191) and
// covariance_matrix (src/utils.f90:745-764) as src/main.f90:524-537 apply it.
//
// The contraction M [n_boot x P] . S [P x n_bin] is small (10^4 x 10^3 x 30 is
// 3e8 FMAs per array, tens of microseconds of FP64 VALU), so it stays on the
// vector units: one wave per sample, lane b owns bin b, the multiplicities are
// wave-uniform scalar loads and the block sums are coalesced loads.  The sums
// run in block order, so the result does not depend on the launch.
//
// Draws: the reference draws with Fortran random_number after an unseeded
// random_seed() (src/main.f90:441), so its multiplicities cannot be
// reproduced; here draw j of sample i is block floor(u P) with u the top 53
// bits of splitmix64(seed + (i << 32) + j) as a fraction, counted with
// integer atomics.  Callers that need given multiplicities pass them.
#include <algorithm>
#include <cmath>

#include "internal.hpp"

namespace tpcf {

void fill_f64(hipStream_t s, double *p, int64_t n);

namespace {

__device__ __forceinline__ uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__global__ void k_zero_i32(int32_t *p, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

// one workgroup per sample: n_draw draws, counted into m[i][.]
__global__ __launch_bounds__(256) void k_boot_draw(int32_t n_block, int64_t n_draw, uint64_t seed,
                                                   int32_t *__restrict__ m)
{
    const int64_t i = blockIdx.x;
    int32_t *row = m + i * (int64_t)n_block;
    for (int64_t j = threadIdx.x; j < n_draw; j += blockDim.x) {
        const uint64_t u = splitmix64(seed + ((uint64_t)i << 32) + (uint64_t)j);
        const double f = (double)(u >> 11) * 0x1.0p-53;
        int32_t p = (int32_t)(f * (double)n_block);
        p = p < n_block ? p : n_block - 1;
        atomicAdd(&row[p], 1);
    }
}

// one wave per sample, lane b = bin b (bins b, b + 64, ... when there are more)
__global__ __launch_bounds__(64) void k_boot_sum(int32_t n_block, int32_t n_bin, const int32_t *__restrict__ m,
                                                 const double *__restrict__ gt, const double *__restrict__ gx,
                                                 const double *__restrict__ k, const double *__restrict__ w,
                                                 double *__restrict__ ot, double *__restrict__ ox,
                                                 double *__restrict__ ow)
{
    const int64_t i = blockIdx.x;
    const int32_t *row = m + i * (int64_t)n_block;
    for (int b0 = 0; b0 < n_bin; b0 += 64) {
        const int b = b0 + (int)threadIdx.x;
        const bool on = b < n_bin;
        double st = 0.0, sx = 0.0, sk = 0.0, sw = 0.0;
        for (int32_t p = 0; p < n_block; ++p) {
            const double mp = (double)row[p];
            if (mp == 0.0) continue;
            if (on) {
                const size_t o = (size_t)p * n_bin + b;
                st = fma(mp, gt[o], st);
                sx = fma(mp, gx[o], sx);
                sk = fma(mp, k[o], sk);
            }
            if (w) sw = fma(mp, w[p], sw);
        }
        if (on) {
            ot[i * n_bin + b] = sk > 0.0 ? st / sk : 0.0;   // src/tpcf.f90:1183-1189
            ox[i * n_bin + b] = sk > 0.0 ? sx / sk : 0.0;
        }
        if (b0 == 0 && threadIdx.x == 0 && ow) ow[i] = sw;
    }
}

// column means of x [n][d]
__global__ void k_col_mean(int32_t n, int32_t d, const double *__restrict__ x, double *__restrict__ mean)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= d) return;
    double s = 0.0;
    for (int32_t i = 0; i < n; ++i) s += x[(size_t)i * d + c];
    mean[c] = s / (double)n;
}

// cov[a][b] = sum_i (x_ia - mean_a)(x_ib - mean_b) / (n - 1), one thread per entry
__global__ void k_cov(int32_t n, int32_t d, const double *__restrict__ x, const double *__restrict__ mean,
                      double *__restrict__ cov)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= d * d) return;
    const int a = e / d, b = e % d;
    double s = 0.0;
    for (int32_t i = 0; i < n; ++i) s = fma(x[(size_t)i * d + a] - mean[a], x[(size_t)i * d + b] - mean[b], s);
    cov[e] = n > 1 ? s / (double)(n - 1) : 0.0;
}

}  // namespace

}  // namespace tpcf

using namespace tpcf;

extern "C" int tpcf_bootstrap(tpcf_ctx *ctx, const tpcf_blocks *blocks, int32_t n_boot, const int32_t *multiplicity,
                              uint64_t seed, double supersampling, uint32_t flags, double *g_t, double *g_x,
                              double *weights, double *cov_t, double *cov_x, int32_t *multiplicity_out)
{
    if (!ctx) return TPCF_ERR_ARG;
    ctx->err.clear();
    if (!blocks || !blocks->g_t || !blocks->g_x || !blocks->k) return fail(ctx, TPCF_ERR_ARG, "blocks is incomplete");
    const int32_t P = blocks->n_block, nb = blocks->n_bin;
    if (P < 1 || nb < 1 || nb > TPCF_MAX_BIN)
        return fail(ctx, TPCF_ERR_ARG, "need n_block >= 1 and 1 <= n_bin <= " + std::to_string(TPCF_MAX_BIN));
    if (n_boot < 1) return fail(ctx, TPCF_ERR_ARG, "n_boot must be >= 1");
    if (!g_t || !g_x) return fail(ctx, TPCF_ERR_ARG, "g_t and g_x outputs are required");
    if (flags & ~(uint32_t)TPCF_DEVICE_INPUTS) return fail(ctx, TPCF_ERR_ARG, "unknown flag");
    const int64_t n_draw = (int64_t)((double)P * supersampling);   // int(n_blocks * supersampling_factor)
    if (!multiplicity && !(n_draw >= 1 && n_draw < ((int64_t)1 << 31)))
        return fail(ctx, TPCF_ERR_ARG, "supersampling must give between 1 and 2^31 draws");
    const bool dev = (flags & TPCF_DEVICE_INPUTS) != 0;
    TPCF_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;

    const int64_t nm = (int64_t)n_boot * P, nbb = (int64_t)n_boot * nb, pb = (int64_t)P * nb;
    // device workspace: m, block sums (3 pb + P), outputs (2 nbb + n_boot), covariances (2 nb^2 + nb)
    const size_t bytes = sizeof(int32_t) * nm + sizeof(double) * (3 * pb + P + 2 * nbb + n_boot + 2 * nb * nb + nb) + 256;
    char *wk;
    TPCF_TRY(ensure(ctx, tpcf_ctx::B_BOOT, bytes, (void **)&wk));
    int32_t *dm = reinterpret_cast<int32_t *>(wk);
    double *dd = reinterpret_cast<double *>(wk + ((sizeof(int32_t) * nm + 15) & ~(size_t)15));
    double *dgt = dd, *dgx = dgt + pb, *dk = dgx + pb, *dw = dk + pb, *ot = dw + P, *ox = ot + nbb, *ow = ox + nbb,
           *ct = ow + n_boot, *cx = ct + (size_t)nb * nb, *mean = cx + (size_t)nb * nb;
    auto in = [&](const void *src, void *dst, size_t n) -> int {
        TPCF_HIP(ctx, hipMemcpyAsync(dst, src, n, dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
        return TPCF_OK;
    };
    TPCF_TRY(in(blocks->g_t, dgt, sizeof(double) * pb));
    TPCF_TRY(in(blocks->g_x, dgx, sizeof(double) * pb));
    TPCF_TRY(in(blocks->k, dk, sizeof(double) * pb));
    if (blocks->weight) TPCF_TRY(in(blocks->weight, dw, sizeof(double) * P));
    else fill_f64(s, dw, P);
    if (multiplicity) {
        TPCF_TRY(in(multiplicity, dm, sizeof(int32_t) * nm));
    } else {
        const unsigned zb = (unsigned)std::min<int64_t>(8192, (nm + 255) / 256);
        hipLaunchKernelGGL(k_zero_i32, dim3(zb), dim3(256), 0, s, dm, nm);
        hipLaunchKernelGGL(k_boot_draw, dim3((unsigned)n_boot), dim3(256), 0, s, P, n_draw, seed, dm);
    }
    hipLaunchKernelGGL(k_boot_sum, dim3((unsigned)n_boot), dim3(64), 0, s, P, nb, dm, dgt, dgx, dk, dw, ot, ox, ow);
    const unsigned cb = (unsigned)((nb * nb + 255) / 256);
    if (cov_t) {
        hipLaunchKernelGGL(k_col_mean, dim3((unsigned)((nb + 63) / 64)), dim3(64), 0, s, n_boot, nb, ot, mean);
        hipLaunchKernelGGL(k_cov, dim3(cb), dim3(256), 0, s, n_boot, nb, ot, mean, ct);
    }
    if (cov_x) {
        hipLaunchKernelGGL(k_col_mean, dim3((unsigned)((nb + 63) / 64)), dim3(64), 0, s, n_boot, nb, ox, mean + 0);
        hipLaunchKernelGGL(k_cov, dim3(cb), dim3(256), 0, s, n_boot, nb, ox, mean + 0, cx);
    }
    TPCF_HIP(ctx, hipGetLastError());
    const hipMemcpyKind ok = dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    TPCF_HIP(ctx, hipMemcpyAsync(g_t, ot, sizeof(double) * nbb, ok, s));
    TPCF_HIP(ctx, hipMemcpyAsync(g_x, ox, sizeof(double) * nbb, ok, s));
    if (weights) TPCF_HIP(ctx, hipMemcpyAsync(weights, ow, sizeof(double) * n_boot, ok, s));
    if (cov_t) TPCF_HIP(ctx, hipMemcpyAsync(cov_t, ct, sizeof(double) * nb * nb, ok, s));
    if (cov_x) TPCF_HIP(ctx, hipMemcpyAsync(cov_x, cx, sizeof(double) * nb * nb, ok, s));
    if (multiplicity_out) TPCF_HIP(ctx, hipMemcpyAsync(multiplicity_out, dm, sizeof(int32_t) * nm, ok, s));
    TPCF_HIP(ctx, hipStreamSynchronize(s));
    return TPCF_OK;
}
