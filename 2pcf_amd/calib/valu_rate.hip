// Issue-rate calibration of single VALU instruction kinds on one GPU (wave64,
// 8 independent dependency chains per lane, 32 waves requested per SIMD), used
// to read the pair kernel's SQ counters: it gives the rate of FP32 FMA, MUL,
// ADD, FMAC, compare + select, FP64 FMA and reciprocal square root, and
// run under rocprofv3 --pmc it shows what SQ_INSTS_VALU and
// SQ_ACTIVE_INST_VALU read for a VALU-saturated kernel.
//
//   hipcc --offload-arch=gfx950 -O3 -Xclang -target-feature -Xclang -packed-fp32-ops \
//         2pcf_amd/calib/valu_rate.hip -o _build/valu_rate && _build/valu_rate
#include <hip/hip_runtime.h>
#include <cstdio>

template <int OP>
__global__ __launch_bounds__(256) void k(float *out, int iters, float a, float b)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    float x[8];
    double d[8];
    for (int c = 0; c < 8; ++c) { x[c] = t * 1e-7f + c; d[c] = x[c]; }
    float y = t * 3e-7f;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            if (OP == 0) x[c] = fmaf(x[c], a, b);          // v_fma_f32 (VOP3)
            if (OP == 1) x[c] = x[c] * a;                  // v_mul_f32
            if (OP == 2) x[c] = x[c] + a;                  // v_add_f32
            if (OP == 3) x[c] = fmaf(y, a, x[c]);          // v_fmac_f32 (VOP2)
            if (OP == 4) x[c] = (x[c] < b) ? x[c] + 0.0f : a;   // cmp + cndmask
            if (OP == 5) d[c] = fma(d[c], (double)a, (double)b);   // v_fma_f64
            if (OP == 6) x[c] = __builtin_amdgcn_rsqf(x[c]);     // v_rsq_f32
        }
        y += 1e-9f;
    }
    float s = 0;
    for (int c = 0; c < 8; ++c) s += x[c] + (float)d[c];
    out[t] = s;
}

// `per`: VALU instructions per chain and iteration (compare + select + add: 3)
template <int OP>
void run(const char *name, float *out, int blocks, int iters, int simds, int per = 1)
{
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0); (void)hipEventCreate(&e1);
    k<OP><<<blocks, 256>>>(out, iters, 0.999f, 1e-3f);
    (void)hipEventRecord(e0);
    k<OP><<<blocks, 256>>>(out, iters, 0.999f, 1e-3f);
    (void)hipEventRecord(e1);
    (void)hipEventSynchronize(e1);
    float ms; (void)hipEventElapsedTime(&ms, e0, e1);
    const double n = blocks * 4.0 * iters * 8 * per;
    printf("%-16s %8.3f ms  %.3e wave-inst/s  %.3f per SIMD per ns\n", name, ms, n / ms * 1e3, n / ms * 1e3 / simds / 1e9);
}

int main()
{
    hipDeviceProp_t p; (void)hipGetDeviceProperties(&p, 0);
    const int blocks = p.multiProcessorCount * 32, iters = 20000;
    const int simds = p.multiProcessorCount * 4;
    float *out; (void)hipMalloc(&out, sizeof(float) * blocks * 256);
    run<0>("fma", out, blocks, iters, simds);
    run<1>("mul", out, blocks, iters, simds);
    run<2>("add", out, blocks, iters, simds);
    run<3>("fmac", out, blocks, iters, simds);
    run<4>("cmp+cndmask+add", out, blocks, iters, simds, 3);
    run<5>("fma_f64", out, blocks, iters, simds);
    run<6>("rsq", out, blocks, iters, simds);
    (void)hipFree(out);
    return 0;
}
