// blocks.hip -- rectangular bootstrap block labels on the device: the
// clustering_scheme rectangular_blocks of survey_marked_point_block_bootstrap
// (src/tpcf.f90:1017-1049), the reference's default bootstrap scheme
// (src/main.f90:100).
//
// Per field the reference lays n_x_blocks x n_y_blocks linear bins over the
// bounding box of the field's objects (create_bins with lin_binning over
// [minval(x), maxval(x)] and [minval(y), maxval(y)], src/utils.f90:139-230)
// and labels each object with find_bin_index of its x and y
// (src/utils.f90:101-137), all in single precision:
//   bin_size = (max - min) / n;  i = int((v - min) / bin_size) + 1,
//   i = n at v == max, -1 outside [min, max] or when i > n;
//   block = (iy - 1) * n_y_blocks + ix + offset     (1-based)
// The row index is multiplied by n_y_blocks, as the reference does: with
// n_x_blocks > n_y_blocks rows share labels, and with n_x_blocks <
// n_y_blocks labels run past the field's n_x_blocks n_y_blocks blocks (into
// the next field's or out of the reference's arrays).  Labels here are the
// reference's minus one (0-based); a label outside [offset, offset +
// n_x_blocks n_y_blocks) is reported as an error instead of being written.
//
// Two launches: per-block partial bounding boxes, then one thread per object
// (each recomputes the two bin sizes with the same single-precision
// operations).  Double inputs are rounded to single precision first, as the
// reference holds positions in single precision.
#include <algorithm>
#include <cmath>
#include <vector>

#include "internal.hpp"

namespace tpcf {

namespace {

constexpr int kRbBlk = 256;

template <typename T>
__global__ __launch_bounds__(kRbBlk) void k_rb_bbox(int64_t n, const T *__restrict__ x, const T *__restrict__ y,
                                                    float *__restrict__ part)
{
    float mnx = INFINITY, mxx = -INFINITY, mny = INFINITY, mxy = -INFINITY, bad = 0.0f;
    for (int64_t i = blockIdx.x * (int64_t)kRbBlk + threadIdx.x; i < n; i += (int64_t)gridDim.x * kRbBlk) {
        const float vx = (float)x[i], vy = (float)y[i];
        if (!isfinite(vx) || !isfinite(vy)) bad = 1.0f;
        mnx = fminf(mnx, vx); mxx = fmaxf(mxx, vx);
        mny = fminf(mny, vy); mxy = fmaxf(mxy, vy);
    }
    __shared__ float sh[5][kRbBlk];
    sh[0][threadIdx.x] = mnx; sh[1][threadIdx.x] = mxx; sh[2][threadIdx.x] = mny; sh[3][threadIdx.x] = mxy;
    sh[4][threadIdx.x] = bad;
    __syncthreads();
    for (int s = kRbBlk / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sh[0][threadIdx.x] = fminf(sh[0][threadIdx.x], sh[0][threadIdx.x + s]);
            sh[1][threadIdx.x] = fmaxf(sh[1][threadIdx.x], sh[1][threadIdx.x + s]);
            sh[2][threadIdx.x] = fminf(sh[2][threadIdx.x], sh[2][threadIdx.x + s]);
            sh[3][threadIdx.x] = fmaxf(sh[3][threadIdx.x], sh[3][threadIdx.x + s]);
            sh[4][threadIdx.x] = fmaxf(sh[4][threadIdx.x], sh[4][threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x < 5) part[5 * blockIdx.x + threadIdx.x] = sh[threadIdx.x][0];
}

// find_bin_index on lin bins of [lo, hi] cut into n (1-based, -1 outside)
__device__ __forceinline__ int lin_bin(float v, float lo, float hi, int n)
{
    if (v < lo || v > hi) return -1;
    if (v == hi) return n;
    const float bin_size = (hi - lo) / (float)n;
    const int i = (int)((v - lo) / bin_size) + 1;
    return i > n ? -1 : i;
}

template <typename T>
__global__ __launch_bounds__(kRbBlk) void k_rb_labels(int64_t n, const T *__restrict__ x, const T *__restrict__ y,
                                                      const float *__restrict__ bb, int32_t nx, int32_t ny,
                                                      int32_t offset, int32_t *__restrict__ lab,
                                                      int32_t *__restrict__ err)
{
    const int64_t i = blockIdx.x * (int64_t)kRbBlk + threadIdx.x;
    if (i >= n) return;
    const int ix = lin_bin((float)x[i], bb[0], bb[1], nx);
    const int iy = lin_bin((float)y[i], bb[2], bb[3], ny);
    const int64_t l = (int64_t)(iy - 1) * ny + ix - 1 + offset;
    if (l < offset || l >= (int64_t)offset + (int64_t)nx * ny) {
        atomicOr(err, 1);
        lab[i] = -1;
    } else {
        lab[i] = (int32_t)l;
    }
}

template <typename T>
int rect_blocks(tpcf_ctx *ctx, int64_t n, const T *x, const T *y, int32_t n_x_blocks, int32_t n_y_blocks,
                int32_t offset, uint32_t flags, int32_t *labels)
{
    if (!ctx) return TPCF_ERR_ARG;
    ctx->err.clear();
    if (!x || !y || !labels) return fail(ctx, TPCF_ERR_ARG, "a required argument is NULL");
    if (n < 1 || n >= ((int64_t)1 << 31)) return fail(ctx, TPCF_ERR_ARG, "n must be in [1, 2^31)");
    if (n_x_blocks < 1 || n_y_blocks < 1 || (int64_t)n_x_blocks * n_y_blocks < 2)
        return fail(ctx, TPCF_ERR_ARG, "at least 2 blocks are required (n_x_blocks, n_y_blocks >= 1)");
    if (offset < 0 || (int64_t)offset + (int64_t)n_x_blocks * n_y_blocks >= ((int64_t)1 << 31))
        return fail(ctx, TPCF_ERR_ARG, "offset must be >= 0 and offset + n_x_blocks n_y_blocks < 2^31");
    if (flags & ~(uint32_t)TPCF_DEVICE_INPUTS) return fail(ctx, TPCF_ERR_ARG, "unknown flag");
    const bool dev = (flags & TPCF_DEVICE_INPUTS) != 0;
    TPCF_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;

    const unsigned nb = (unsigned)((n + kRbBlk - 1) / kRbBlk);
    const unsigned nbb = std::min<unsigned>(nb, 1024);
    const size_t in_bytes = dev ? 0 : ((sizeof(T) * 2 * (size_t)n + 15) & ~(size_t)15);
    const size_t lab_bytes = dev ? 0 : ((sizeof(int32_t) * (size_t)n + 15) & ~(size_t)15);
    const size_t bytes = in_bytes + lab_bytes + sizeof(float) * 5 * nbb + 64;
    char *wk;
    TPCF_TRY(ensure(ctx, tpcf_ctx::B_KMEANS, bytes, (void **)&wk));
    const T *dx = x, *dy = y;
    int32_t *lab = labels;
    char *p = wk;
    if (!dev) {
        T *t = reinterpret_cast<T *>(p);
        TPCF_HIP(ctx, hipMemcpyAsync(t, x, sizeof(T) * n, hipMemcpyHostToDevice, s));
        TPCF_HIP(ctx, hipMemcpyAsync(t + n, y, sizeof(T) * n, hipMemcpyHostToDevice, s));
        dx = t;
        dy = t + n;
        p += in_bytes;
        lab = reinterpret_cast<int32_t *>(p);
        p += lab_bytes;
    }
    float *part = reinterpret_cast<float *>(p);
    int32_t *err = reinterpret_cast<int32_t *>(part + 5 * nbb);
    hipLaunchKernelGGL(k_rb_bbox<T>, dim3(nbb), dim3(kRbBlk), 0, s, n, dx, dy, part);
    std::vector<float> hp(5 * nbb);
    TPCF_HIP(ctx, hipMemcpyAsync(hp.data(), part, sizeof(float) * hp.size(), hipMemcpyDeviceToHost, s));
    TPCF_HIP(ctx, hipStreamSynchronize(s));
    float bb[4] = {INFINITY, -INFINITY, INFINITY, -INFINITY};
    bool bad = false;
    for (unsigned b = 0; b < nbb; ++b) {
        bb[0] = std::min(bb[0], hp[5 * b]);
        bb[1] = std::max(bb[1], hp[5 * b + 1]);
        bb[2] = std::min(bb[2], hp[5 * b + 2]);
        bb[3] = std::max(bb[3], hp[5 * b + 3]);
        bad = bad || hp[5 * b + 4] != 0.0f;
    }
    if (bad) return fail(ctx, TPCF_ERR_NONFINITE, "non-finite coordinate");
    // the field's box (the first 4 floats of the partials) and the error flag
    TPCF_HIP(ctx, hipMemcpyAsync(part, bb, sizeof bb, hipMemcpyHostToDevice, s));
    TPCF_HIP(ctx, hipMemsetAsync(err, 0, sizeof(int32_t), s));
    hipLaunchKernelGGL(k_rb_labels<T>, dim3(nb), dim3(kRbBlk), 0, s, n, dx, dy, part, n_x_blocks, n_y_blocks, offset,
                       lab, err);
    int32_t herr = 0;
    TPCF_HIP(ctx, hipMemcpyAsync(&herr, err, sizeof herr, hipMemcpyDeviceToHost, s));
    if (!dev) TPCF_HIP(ctx, hipMemcpyAsync(labels, lab, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
    TPCF_HIP(ctx, hipStreamSynchronize(s));
    TPCF_HIP(ctx, hipGetLastError());
    if (herr)
        return fail(ctx, TPCF_ERR_ARG, "a block label leaves [offset, offset + n_x_blocks n_y_blocks): the "
                                       "reference's row stride n_y_blocks exceeds n_x_blocks there "
                                       "(n_x_blocks < n_y_blocks), or an object fell outside its field's bins");
    return TPCF_OK;
}

}  // namespace

}  // namespace tpcf

using namespace tpcf;

extern "C" int tpcf_rect_blocks_f32(tpcf_ctx *ctx, int64_t n, const float *x, const float *y, int32_t n_x_blocks,
                                    int32_t n_y_blocks, int32_t offset, uint32_t flags, int32_t *labels)
{
    return rect_blocks<float>(ctx, n, x, y, n_x_blocks, n_y_blocks, offset, flags, labels);
}

extern "C" int tpcf_rect_blocks_f64(tpcf_ctx *ctx, int64_t n, const double *x, const double *y, int32_t n_x_blocks,
                                    int32_t n_y_blocks, int32_t offset, uint32_t flags, int32_t *labels)
{
    return rect_blocks<double>(ctx, n, x, y, n_x_blocks, n_y_blocks, offset, flags, labels);
}
