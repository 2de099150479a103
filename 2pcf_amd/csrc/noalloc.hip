// noalloc.hip -- tpcf_shear_f32_noalloc: tpcf_shear_f32 without device
// allocation, for callers that cannot call hipMalloc (serving loops that must
// not allocate, harnesses that forbid it).
//
// Every device buffer of the call (sort keys and their scratch, source
// records, segments, chunks, work items, accumulators, marks, staging) is
// carved from one file-scope device array, `scratch`, of kScratchBytes.  Each
// GPU the library runs on holds its own copy of the array until the process
// ends; its device address is looked up once per device with
// hipGetSymbolAddress (the array's name in host code is only a host-side
// stand-in) and that address is what the kernels receive.  A call whose
// buffers do not fit returns TPCF_ERR_ALLOC with a message naming the limit.
//
// The array keeps the last call's contents: the call clears what must start
// at zero (counters, accumulators, error flags) with fill kernels, as the
// allocating entry point does.  Calls that overlap on one GPU (two host
// threads, or two contexts on different streams) share the array and race:
// they must be serialised by the caller.
#include <mutex>

#include "internal.hpp"

namespace tpcf {

constexpr size_t kScratchBytes = size_t(256) << 20;   // 256 MiB per GPU

static __device__ __align__(256) char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;

std::mutex g_scratch_mutex;
void *g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on ctx's device (looked up once per device,
// with that device current).
int scratch_base(tpcf_ctx *ctx, char **base)
{
    const int dev = ctx->device;
    if (dev < 0 || dev >= kMaxDevices) return fail(ctx, TPCF_ERR_ARG, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        TPCF_HIP(ctx, hipSetDevice(dev));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(ctx, TPCF_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *base = static_cast<char *>(g_scratch_ptr[dev]);
    return TPCF_OK;
}

// A lower bound of the device bytes a shear call needs (the arena enforces
// the exact figure): sort keys, indices and records of both catalogues,
// uploads of host inputs, and the patch-pair accumulators with their
// reduction.
size_t lower_bound_bytes(int64_t n_a, int64_t n_b, int32_t n_pa, int32_t n_pb, int32_t n_bin, bool host_inputs)
{
    const size_t pp_total = (size_t)n_pa * (size_t)n_pb * (size_t)n_bin;
    size_t b = (size_t)n_b * (2 * 4 + 2 * 4 + sizeof(SrcRec<float>) + 4) + (size_t)n_a * (2 * 8 + 2 * 4 + 5 * 4);
    b += pp_total * ((size_t)kSlots * (kNQ * 8 + 8) + (kNQ + 1) * 8);
    if (host_inputs) b += (size_t)n_a * (5 * 4 + 4) + (size_t)n_b * (7 * 4 + 4);
    return b;
}

}  // namespace

}  // namespace tpcf

using namespace tpcf;

extern "C" int tpcf_shear_f32_noalloc(tpcf_ctx *ctx, int64_t n_A, const float *x_A, const float *y_A,
                                      const float *z_A, const float *val_A, const float *w_A, int64_t n_B,
                                      const float *x_B, const float *y_B, const float *z_B, const float *e1_B,
                                      const float *e2_B, const float *w_B, const float *m_B, const tpcf_bins *bins,
                                      uint32_t flags, const int32_t *patch_A, const int32_t *patch_B,
                                      int32_t n_patch_A, int32_t n_patch_B, tpcf_out *out)
{
    if (!ctx) return TPCF_ERR_ARG;
    ctx->err.clear();
    TPCF_TRY(check_bins(ctx, bins));
    if (n_A < 0 || n_B < 0) return fail(ctx, TPCF_ERR_ARG, "catalogue sizes must be >= 0");
    const int32_t pa = patch_A ? n_patch_A : 1, pb = patch_B ? n_patch_B : 1;
    if (pa < 1 || pb < 1) return fail(ctx, TPCF_ERR_ARG, "n_patch must be >= 1");
    const size_t need = lower_bound_bytes(n_A, n_B, pa, pb, bins->n_bin, !(flags & TPCF_DEVICE_INPUTS));
    if (need > kScratchBytes)
        return fail(ctx, TPCF_ERR_ALLOC, "the call needs at least " + std::to_string(need >> 20) +
                                             " MiB of device memory, more than the " +
                                             std::to_string(kScratchBytes >> 20) +
                                             " MiB scratch region of the allocation-free entry point");
    char *base = nullptr;
    TPCF_TRY(scratch_base(ctx, &base));
    // an empty buffer table for this call: every ensure() carves from the
    // scratch region; the context's own buffers are put back afterwards
    DevBuf saved[tpcf_ctx::B_COUNT];
    for (int i = 0; i < tpcf_ctx::B_COUNT; ++i) {
        saved[i] = ctx->buf[i];
        ctx->buf[i] = DevBuf{};
    }
    ctx->arena = base;
    ctx->arena_cap = kScratchBytes;
    ctx->arena_off = 0;
    const int rc = correlate<float>(ctx, SHEAR, n_A, x_A, y_A, z_A, val_A, w_A, n_B, x_B, y_B, z_B, e1_B, e2_B, w_B,
                                    m_B, bins, flags, patch_A, patch_B, n_patch_A, n_patch_B, out);
    // a failed call may leave work queued on the region: it ends before the
    // next call can reuse the region
    if (rc) (void)hipStreamSynchronize(ctx->stream);
    for (int i = 0; i < tpcf_ctx::B_COUNT; ++i) ctx->buf[i] = saved[i];
    ctx->arena = nullptr;
    ctx->arena_cap = ctx->arena_off = 0;
    return rc;
}
