// internal.hpp -- shared records, device helpers and host plumbing of the
// gfx950 pair-counting library.  Not part of the public C-ABI (tpcf_gpu.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "tpcf_gpu.h"

namespace tpcf {

constexpr int kWave = 64;          // CDNA wavefront width
constexpr int kWavesPerBlock = 4;  // pair kernel: 4 independent waves per workgroup
constexpr int kChunk = 64;         // lenses per chunk = one lens per lane
constexpr int kSlots = 8;          // per-XCD copies of the patch-pair accumulators
constexpr int kNQ = 5;             // accumulated FP64 quantities: Theta, G1, G2, W, K
constexpr uint32_t kSegMax = 256;  // longest source segment (bounds per-segment FP32 sums)
constexpr int kRecPad = 4;         // readable records past the last source (prefetch slack)
// Bins one pair-kernel pass handles: one squared edge per wave lane (edges
// 0..63) and one bin per lane in the per-bin registers.  A call with more
// bins runs the pass once per group of up to kGroupBins consecutive bins.
constexpr int kGroupBins = 63;

inline int bin_groups(int n_bin) { return (n_bin + kGroupBins - 1) / kGroupBins; }

enum Kind { SHEAR = 0, KAPPA = 1 };

// One source, sorted into (patch, cell) order.  Eight values of the input
// precision: x, y, w, w*(1+m), s1*w, s2*w, 2*s1*w, 2*s2*w where (s1,s2) are
// (e1,e2) for shear and (kappa_E,kappa_B) for kappa (kappa uses w*(1+m) = w).
template <typename T>
struct alignas(sizeof(T) * 8) SrcRec {
    T x, y, w, wk, s1w, s2w, s1w2, s2w2;
};

// A non-empty (source patch, cell) run of sorted sources with its tight
// bounding box and the sums used when the whole run falls into one bin.
template <typename T>
struct alignas(16) Seg {
    uint32_t start, count, pb, pad;
    T x0, x1, y0, y1;
    T zmin, zpad;
    double W, K, S1, S2;
};

// Up to 64 consecutive sorted lenses of one lens patch.
template <typename T>
struct alignas(16) Chunk {
    uint32_t start, count, pa, pad;
    T bx0, bx1, by0, by1;
    T zmax, zpad;
    double wsum, asum;
};

// Uniform cell grid laid over the sources.  Cell of a coordinate:
// clamp(floor((x - x0) * inv), 0, n - 1) in FP64 -- a monotone function, so
// conservative coordinate bounds map to conservative cell bounds.
struct Grid {
    double x0, y0, cell, inv;
    int32_t nx, ny;
    int64_t ncells;
};

__host__ __device__ inline int32_t cell_of(double v, double v0, double inv, int32_t n)
{
    double f = floor((v - v0) * inv);
    if (!(f >= 0.0)) return 0;   // also catches NaN
    if (f >= (double)(n - 1)) return n - 1;
    return (int32_t)f;
}

// Cell rows a lens box [bx0,bx1]x[by0,by1] can reach within x_max (plus the
// slack mg), before clipping to a patch.
__host__ __device__ inline void row_span(const Grid &g, double by0, double by1, double xm, double mg, int32_t &cy0,
                                         int32_t &cy1)
{
    cy0 = cell_of(by0 - xm - mg, g.y0, g.inv, g.ny);
    cy1 = cell_of(by1 + xm + mg, g.y0, g.inv, g.ny);
}

// Cell columns of row cy that can hold a source within x_max of the box;
// false when the whole row is out of reach.  Conservative: every source in
// reach lies in [cx0, cx1] because cell_of is monotone.
__host__ __device__ inline bool col_span(const Grid &g, double bx0, double bx1, double by0, double by1, int32_t cy,
                                         double xm, double mg, int32_t &cx0, int32_t &cx1)
{
    const double ry0 = g.y0 + (double)cy * g.cell - mg, ry1 = g.y0 + (double)(cy + 1) * g.cell + mg;
    const double dyl = fmax(0.0, fmax(ry0 - by1, by0 - ry1));
    if (dyl > xm) return false;
    const double hw = sqrt(fmax(xm * xm - dyl * dyl, 0.0)) + mg;
    cx0 = cell_of(bx0 - hw, g.x0, g.inv, g.nx);
    cx1 = cell_of(bx1 + hw, g.x0, g.inv, g.nx);
    return true;
}

// Hilbert-curve index of a point on a 65536 x 65536 lattice (the classic
// iterative quadrant-rotation construction): consecutive indices are
// neighbouring lattice points, so runs of the sorted order are compact.
__host__ __device__ inline uint32_t hilbert16(uint32_t x, uint32_t y)
{
    uint32_t d = 0;
    for (uint32_t s = 1u << 15; s > 0; s >>= 1) {
        const uint32_t rx = (x & s) ? 1u : 0u, ry = (y & s) ? 1u : 0u;
        d += s * s * ((3u * rx) ^ ry);
        if (ry == 0) {
            if (rx == 1) {
                x = 0xffffu - x;
                y = 0xffffu - y;
            }
            const uint32_t t = x;
            x = y;
            y = t;
        }
    }
    return d;
}

// One unit of pair-kernel work: the source cell rows row0..row1 of a lens
// chunk (one band of rows, clipped to the rows the chunk can reach).  Items
// with chunk == kNoChunk pad the per-XCD item ranges to equal length.
struct Item {
    uint32_t chunk;
    uint16_t row0, row1;
};
constexpr uint32_t kNoChunk = 0xffffffffu;

// ---- wavefront reductions (DPP row shifts + row broadcasts, gfx9 family) ----
template <int CTRL, int ROWMASK = 0xf, bool BC = true>
__device__ inline int dpp_i(int v)
{
    return __builtin_amdgcn_update_dpp(0, v, CTRL, ROWMASK, 0xf, BC);
}

// Sum over all 64 lanes; every lane must be active.  Result is uniform.
__device__ inline float wave_sum(float v)
{
    v += __int_as_float(dpp_i<0x111>(__float_as_int(v)));
    v += __int_as_float(dpp_i<0x112>(__float_as_int(v)));
    v += __int_as_float(dpp_i<0x114>(__float_as_int(v)));
    v += __int_as_float(dpp_i<0x118>(__float_as_int(v)));
    v += __int_as_float(dpp_i<0x142, 0xa, false>(__float_as_int(v)));
    v += __int_as_float(dpp_i<0x143, 0xc, false>(__float_as_int(v)));
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

template <int CTRL, int ROWMASK = 0xf, bool BC = true>
__device__ inline double dpp_d(double v)
{
    int2 p = *reinterpret_cast<int2 *>(&v);
    p.x = dpp_i<CTRL, ROWMASK, BC>(p.x);
    p.y = dpp_i<CTRL, ROWMASK, BC>(p.y);
    return *reinterpret_cast<double *>(&p);
}

__device__ inline double wave_sum(double v)
{
    v += dpp_d<0x111>(v);
    v += dpp_d<0x112>(v);
    v += dpp_d<0x114>(v);
    v += dpp_d<0x118>(v);
    v += dpp_d<0x142, 0xa, false>(v);
    v += dpp_d<0x143, 0xc, false>(v);
    int2 p = *reinterpret_cast<int2 *>(&v);
    p.x = __builtin_amdgcn_readlane(p.x, 63);
    p.y = __builtin_amdgcn_readlane(p.y, 63);
    return *reinterpret_cast<double *>(&p);
}

// Sum over all 64 lanes of an integer; every lane must be active.
__device__ inline uint32_t wave_sum_u32(uint32_t v)
{
    int x = (int)v;
    x += dpp_i<0x111>(x);
    x += dpp_i<0x112>(x);
    x += dpp_i<0x114>(x);
    x += dpp_i<0x118>(x);
    x += dpp_i<0x142, 0xa, false>(x);
    x += dpp_i<0x143, 0xc, false>(x);
    return (uint32_t)__builtin_amdgcn_readlane(x, 63);
}

// ---- host plumbing ----
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
};

struct ProbeCounters {
    unsigned long long seg[3];
    unsigned long long pairs[3];
};

}  // namespace tpcf

struct tpcf_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t own_stream = nullptr;
    hipEvent_t ev[6] = {};
    double opt_cell = 0.0;
    // 0: the automatic cell size uses the precision's own mean sources per
    // cell (cell_target below).  Work-item bands: 0.12-0.45 x_max are within
    // 0.6 % of each other on the 20M x 20M clustered shear workload, and 0.2
    // is the best on the FP64 1M x 2M one (0.3: 1.3 % slower, 0.45: 4 %)
    double opt_cell_target = 0.0;
    double opt_split = 0.2;       // work-item band height in units of x_max
    bool opt_counters = false;
    tpcf_stats stats{};
    std::string err;
    // Allocation-free calls: while `arena` is set, ensure() carves the named
    // buffers out of [arena, arena + arena_cap) with a bump pointer instead of
    // calling hipMalloc (the caller swaps in an empty buffer table first).
    char *arena = nullptr;
    size_t arena_cap = 0, arena_off = 0;
    // named device buffers, grown on demand and reused between calls
    enum {
        B_IN0, B_IN1, B_IN2, B_IN3, B_IN4, B_IN5, B_IN6, B_IN7, B_IN8, B_IN9, B_IN10, B_IN11,
        B_PATCH_A, B_PATCH_B,
        B_PARTIAL, B_ERR,
        B_SKEY, B_SKEY2, B_SIDX, B_SIDX2, B_COUNTS, B_CSTART, B_SEGFIRST, B_SREC, B_SZ, B_SEGS, B_PBOX,
        B_LKEY, B_LKEY2, B_LIDX, B_LIDX2, B_PCOUNT,
        B_LX, B_LY, B_LZ, B_LA, B_LW, B_CHUNKS,
        B_PP, B_PPN, B_MARKS, B_OUT, B_TMP, B_PROBE, B_E2,
        B_BLK, B_BLKCH, B_WORK, B_ITEMS,
        B_GRAD_A, B_GRAD_B, B_ONES, B_ZEROS,
        // spherical geometry: FP64 inputs, projected coordinates, lens frames, sorted extras
        B_SPH_IN, B_SPH_LP, B_SPH_LF, B_SPH_SEGP, B_SPH_CHP, B_SPH_PART, B_SPH_RF,
        B_BOOT, B_KMEANS,
        B_COUNT
    };
    tpcf::DevBuf buf[B_COUNT];
};

// Mean sources per cell of the automatic cell size: TPCF_OPT_CELL_TARGET when
// set, else 128 for FP32 geometry and 64 for FP64.  Measured on one GPU: the
// 20M x 20M clustered shear workload (FP32) runs 2.7 % faster at 112-160
// than at 64 (192: 0.8 % slower than 128), the 5M x 5M adjoint (FP32) 3 %
// faster at 128; the FP64 1M x 2M workload is best at 64-96 (128: 1 %
// slower, 192: 6 %), the FP64 kappa-kappa 100k call at 64 (128: 5 %), and
// the sphere (FP64) within 0.4 % from 64 to 160.  The FP64 pair loops cost
// more per pair against the same per-visit work, which moves the balance
// between visits and edge pairs to smaller cells.
template <typename T>
inline double cell_target(const tpcf_ctx *ctx)
{
    return ctx->opt_cell_target > 0.0 ? ctx->opt_cell_target : (sizeof(T) == 4 ? 128.0 : 64.0);
}

namespace tpcf {

int fail(tpcf_ctx *ctx, int code, const std::string &msg);
int ensure(tpcf_ctx *ctx, int which, size_t bytes, void **ptr);
// TPCF_ERR_BINS unless 1 <= n_bin <= TPCF_MAX_BIN and the edges are finite,
// non-negative and increasing (what the pair kernel's bin search assumes).
int check_bins(tpcf_ctx *ctx, const tpcf_bins *b);

#define TPCF_HIP(ctx, call)                                                              \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            return ::tpcf::fail((ctx), TPCF_ERR_HIP,                                     \
                                std::string(#call) + ": " + hipGetErrorString(e_));      \
    } while (0)

#define TPCF_TRY(expr)          \
    do {                        \
        int rc_ = (expr);       \
        if (rc_) return rc_;    \
    } while (0)

// Arguments of one correlation call after validation; pointers are device
// pointers of the input precision T.
template <typename T>
struct CallArgs {
    int kind;
    int64_t n_a, n_b;
    const T *xa, *ya, *za, *vala, *wa;
    const T *xb, *yb, *zb, *s1b, *s2b, *wb, *mb;   // zb, mb NULL for kappa
    const int32_t *patch_a, *patch_b;              // device, nullable
    int32_t n_pa, n_pb;
    const tpcf_bins *bins;
    bool marks;
    bool verbose;
    int zsign = 1;              // z as stored for the z test: +z, or -z to reverse z_A > z_B
    bool marks_device = false;  // keep the marks on the device (adjoint) instead of copying them out
    // Spherical geometry (T = double only): xa/ya and xb/yb then hold the
    // projected grid coordinates (q, s) of the unit vectors, sph_pa/sph_pb the
    // third rotated component p, sph_frame the lens' east and north unit
    // vectors in the rotated basis ([n_a][6]); kbins has the chord edges
    // 2 sin(e/2) that the neighbour search and the pair kernel use, while bins
    // (the angles) give the outputs.  kbins == bins on the flat sky.
    const tpcf_bins *kbins = nullptr;
    const double *sph_pa = nullptr, *sph_pb = nullptr, *sph_frame = nullptr;
    // Shear on the sphere with the catalogues' roles swapped (the adjoint's
    // source-side pass): sph_frame then holds (u x P, P) of the lane objects,
    // P the pole in the rotated basis, and the angle is taken at the streamed
    // object.
    bool sph_rev = false;
    double sph_pole[3] = {0.0, 0.0, 1.0};   // the pole (latitude pi/2) in the rotated basis
};

// Device-side result of the preparation stage.
template <typename T>
struct Prepared {
    Grid grid;
    int64_t n_src;         // sources kept
    int64_t n_lens;        // lenses kept
    int64_t n_chunks;
    int64_t n_segs;
    int64_t key_space;     // n_pb * ncells
    const SrcRec<T> *srec;
    const T *sz;           // sorted source z (shear)
    const Seg<T> *segs;
    const uint32_t *seg_first;   // [key_space + 1]
    const int4 *pbox;            // per source patch: cx0, cx1, cy0, cy1 (cx0 > cx1 when empty)
    const Chunk<T> *chunks;
    const T *lx, *ly, *lz, *la, *lw;
    const uint32_t *lidx;        // sorted lens -> original index
    const uint32_t *cstart;      // [key_space + 1] first sorted source of each key
    const Item *items;           // work items of the pair kernel
    int64_t n_items;
    double margin;               // slack of the cell-range bounds
    bool m_zero;                 // shear: every source has m == 0 (then K == W)
    // what refresh_values needs to re-read the values on the same structure
    const uint32_t *sidx;        // sorted source -> original index
    const uint32_t *counts;      // [key_space + 1] sources per key
    int32_t *err;                // device flags
    // spherical geometry (else NULL): sorted lens p and frames [n_lens][6]
    // (east, north), per segment and per chunk (min, max) of p
    const double *sph_lp, *sph_lf;
    const double2 *sph_segp, *sph_chp;
    bool sph_rev;                // angles at the streamed object (CallArgs::sph_rev)
};

template <typename T>
int prepare(tpcf_ctx *ctx, const CallArgs<T> &a, Prepared<T> &p);

// Re-reads the values (val, w, z, s1, s2, m) of the same catalogues into the
// prepared structure on stream s: source records and segment sums, sorted
// lens values and chunk sums.  Kernel launches only (graph-capturable).  The
// positions and the masks (w_A == 0, z < 0) are those of the preparation: a
// kept source whose new z is < 0 sets bit 8 of p.err, a kept lens whose new w
// is 0 (or, for shear, z < 0) sets bit 16 (p.err is cleared first).
template <typename T>
int refresh_values(tpcf_ctx *ctx, const CallArgs<T> &a, const Prepared<T> &p, hipStream_t s);

// Arguments of the pair kernels shared by every catalogue (binning, layout).
template <typename T>
struct PairArgs {
    int64_t n_items;
    int64_t per_xcd;       // blocks per XCD of the remapped grid
    int32_t n_bin, counters;   // n_bin: the bins of this pass (a group of the call's bins)
    int32_t pp_stride;     // bins per row of the accumulators (all bins of the call)
    double x_max;          // outer edge of this pass (FP64)
    int32_t xmin_zero;     // x_min == 0: pairs at zero separation fall into bin 0
    T cls_lo, cls_hi;      // classification margins on squared box distances
    float band;            // relative ambiguity band on d^2 (FP32 path)
    float e2f[kGroupBins + 1];   // squared edges rounded to FP32 (+inf past n_bin)
    double e2d[kGroupBins + 1];  // squared edges in FP64; for the last group e2d[n_bin] = next double above x_max^2
};

// One catalogue's prepared state, as the pair kernels read it.
// Per-lens marks on the device: component planes [n_bin][4][n_lens] over the
// sorted lenses (q = xi_E or P, xi_B or Q, K, W), so the lenses of one wave,
// which are consecutive, add to consecutive addresses.
__host__ __device__ inline size_t mark_at(int64_t n_lens, int bin, int q, int64_t li)
{
    return ((size_t)bin * 4 + (size_t)q) * (size_t)n_lens + (size_t)li;
}

template <typename T>
struct CatArgs {
    Grid g;
    int32_t n_pa, n_pb;
    double margin;                  // absolute slack for cell-range bounds
    const Chunk<T> *chunks;
    const T *lx, *ly, *lz, *la, *lw;
    const Seg<T> *segs;
    const uint32_t *seg_first;
    const int4 *pbox;
    const SrcRec<T> *srec;
    const T *sz;
    double *pp;                     // [slot][P_A][P_B][n_bin][5]
    unsigned long long *ppn;        // [slot][P_A][P_B][n_bin]
    double *marks;                  // [n_bin][4][n_lens] (mark_at) or NULL
    int64_t marks_n;                // n_lens, the stride of the marks planes
    const double *sph_lp, *sph_lf;  // spherical geometry (else NULL)
    const double2 *sph_segp, *sph_chp;
    bool sph_rev;                   // angles at the streamed object
};



template <typename T>
int make_pair_args(tpcf_ctx *ctx, const tpcf_bins *bins, int group, PairArgs<T> &A, const T **e2lane_out);
template <typename T>
CatArgs<T> cat_args(const Prepared<T> &p, int32_t n_pa, int32_t n_pb, double *pp, unsigned long long *ppn,
                    double *marks);
template <typename T>
int launch_pairs(tpcf_ctx *ctx, hipStream_t s, int kind, bool m_zero, bool marks, const PairArgs<T> &A,
                 const CatArgs<T> &C, const T *e2lane, const Item *items, unsigned long long *probe);
// Per-patch-pair and per-bin reduction of the accumulators (kernels on s):
// ppo [5][P_B][P_A][n_bin], ppno [P_B][P_A][n_bin], tot [n_bin][5] FP64 sums
// followed by [n_bin] int64 pair counts.
int launch_reduce(tpcf_ctx *ctx, hipStream_t s, int32_t n_pa, int32_t n_pb, int32_t n_bin, const double *pp,
                  const unsigned long long *ppn, double *ppo, long long *ppno, double *tot);
void fill_f64(hipStream_t s, double *p, int64_t n);
void fill_u64(hipStream_t s, unsigned long long *p, int64_t n);

// A whole correlation call: argument checks, uploads (unless
// TPCF_DEVICE_INPUTS), then run_call or the spherical path (api.hip).
template <typename T>
int correlate(tpcf_ctx *ctx, int kind, int64_t n_a, const T *xa, const T *ya, const T *za, const T *vala,
              const T *wa, int64_t n_b, const T *xb, const T *yb, const T *zb, const T *s1b, const T *s2b,
              const T *wb, const T *mb, const tpcf_bins *bins, uint32_t flags, const int32_t *pat_a,
              const int32_t *pat_b, int32_t n_pa, int32_t n_pb, tpcf_out *out);

// The correlation call once the inputs are on the device (api.hip).
template <typename T>
int run_call(tpcf_ctx *ctx, CallArgs<T> &a, uint32_t flags, tpcf_out *out, double t0, Prepared<T> *pout);

// Spherical coordinates (sphere.hip): projects the device inputs, converts
// them to FP64 and runs the FP64 path with the chord bins.
template <typename T>
int sph_correlate(tpcf_ctx *ctx, const CallArgs<T> &a, uint32_t flags, tpcf_out *out, double t0);
// After prepare: the sorted lens p and frames, the records' p and the
// per-segment / per-chunk p ranges.
int sph_finish(tpcf_ctx *ctx, const CallArgs<double> &a, Prepared<double> &p);
// The chord edges 2 sin(e/2) of angle bins (what the spherical neighbour
// search and pair kernel compare with).
int sph_chord_bins(tpcf_ctx *ctx, const tpcf_bins *bins, tpcf_bins &kb);
// One FP32 -> FP64 copy of an input array the spherical path made.
struct F64Copy {
    const float *src;
    double *dst;
    int64_t n;
};
// The front half of sph_correlate: the basis, the projected positions and
// frames, and FP64 copies of the values (listed in *copies when given, so a
// batch can repeat them on new values), into d with d.kbins = kb.
template <typename T>
int sph_setup(tpcf_ctx *ctx, const CallArgs<T> &a, const tpcf_bins *kb, CallArgs<double> &d,
              std::vector<F64Copy> *copies);
void launch_to_f64(hipStream_t s, const float *in, double *out, int64_t n);
// Frames of the sources of d for a pass with the roles swapped (CallArgs::
// sph_rev): (u_B x P, P) per source [n_b][6] in the rotated basis, into the
// context's buffer B_SPH_RF.
int sph_rev_frames(tpcf_ctx *ctx, const CallArgs<double> &d, const double **frames);
// The sources' p into record slot s1w2 (after a refresh_values re-gather).
void sph_records_p(hipStream_t s, const Prepared<double> &p, const double *pb);

// The per-lens marks of a pass (FP64 planes, sorted lens order) into the
// caller's device arrays out[0..4] = xi_E, xi_B, K, w marks [n_a][n_bin] and
// marks_weight [n_a] of rb bytes per element (any may be NULL): the K planes
// are first copied from the W planes when k_from_w, every output is zeroed,
// then the pair set's rows are written.  Kernel launches only.
int launch_marks_out(tpcf_ctx *ctx, hipStream_t s, int64_t n_lens, int32_t n_bin, int64_t n_a, const uint32_t *lidx,
                     double *mk, void *const out[5], int rb, bool k_from_w);

double now_ms();

template <typename T>
int run_pairs(tpcf_ctx *ctx, const CallArgs<T> &a, const Prepared<T> &p, double *pp, unsigned long long *ppn,
              double *marks);

}  // namespace tpcf
