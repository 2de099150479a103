// prepare.hip -- device-side replacement of the reference's kd-tree build
// (grow_tree, src/kdtree.f90:73-205, called at src/tpcf.f90:255-257).
//
// Instead of a pointer tree the catalogues are sorted on the device:
//  * sources by key (patch_B, cell) on a uniform grid: every non-empty
//    (patch, cell) run becomes a "segment" with a tight bounding box and the
//    sums needed when all its pairs fall into one bin;
//  * lenses by (patch_A, Hilbert index) and cut into compact chunks of up to
//    64 lenses (one lens per lane), each with its bounding box; a chunk's
//    work is cut into items of one band of source cell rows each, ordered
//    band by band on each XCD, so that the waves running together read the
//    same rows of sources (from L2), and each XCD gets an equal estimated
//    share of the work.
// Sources masked by the reference (z_B < 0, src/tpcf.f90:306-309) and lenses
// it skips (w_A == 0, src/tpcf.f90:285 / :474; z_A < 0 for shear) are dropped
// here, which removes the same pairs.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <climits>
#include <cmath>

#include "internal.hpp"

namespace tpcf {

namespace {

constexpr int kBlk = 256;

template <typename U>
__global__ void k_fill(U *p, U v, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

template <typename U>
void launch_fill(hipStream_t s, U *p, U v, int64_t n)
{
    if (n <= 0) return;
    int64_t blocks = std::min<int64_t>((n + kBlk - 1) / kBlk, 8192);
    hipLaunchKernelGGL(k_fill<U>, dim3((unsigned)blocks), dim3(kBlk), 0, s, p, v, n);
}

// counter[key] += 1 for every active lane with add set.  When all of them
// share one key (a single patch, or neighbouring objects) the wave adds its
// total with one atomic instead of up to 64 on one address.
__device__ __forceinline__ void wave_count(uint32_t *counter, uint32_t key, bool add)
{
    const uint64_t act = __ballot(add);
    if (!act) return;
    const int lead = __ffsll((unsigned long long)act) - 1;
    const uint32_t k0 = __shfl(key, lead);
    if (__ballot(add && key != k0) == 0) {
        if ((int)__lane_id() == lead) atomicAdd(&counter[k0], (uint32_t)__popcll(act));
    } else if (add) {
        atomicAdd(&counter[key], 1u);
    }
}

// *flags |= bit if any active lane has set: one atomic per wave.
__device__ __forceinline__ void wave_flag(int32_t *flags, int32_t bit, bool set)
{
    const uint64_t m = __ballot(set);
    if (m && (int)__lane_id() == __ffsll((unsigned long long)m) - 1) atomicOr(flags, bit);
}

// Per-block min/max of x and y and count of non-finite coordinates.
template <typename T>
__global__ __launch_bounds__(kBlk) void k_bbox(int64_t n, const T *__restrict__ x, const T *__restrict__ y,
                                               double *__restrict__ out)
{
    double mnx = INFINITY, mxx = -INFINITY, mny = INFINITY, mxy = -INFINITY, bad = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double vx = (double)x[i], vy = (double)y[i];
        if (!isfinite(vx) || !isfinite(vy)) {
            bad += 1.0;
            continue;
        }
        mnx = fmin(mnx, vx); mxx = fmax(mxx, vx);
        mny = fmin(mny, vy); mxy = fmax(mxy, vy);
    }
    __shared__ double sh[5][kBlk];
    sh[0][threadIdx.x] = mnx; sh[1][threadIdx.x] = mxx; sh[2][threadIdx.x] = mny; sh[3][threadIdx.x] = mxy;
    sh[4][threadIdx.x] = bad;
    __syncthreads();
    for (int s = kBlk / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sh[0][threadIdx.x] = fmin(sh[0][threadIdx.x], sh[0][threadIdx.x + s]);
            sh[1][threadIdx.x] = fmax(sh[1][threadIdx.x], sh[1][threadIdx.x + s]);
            sh[2][threadIdx.x] = fmin(sh[2][threadIdx.x], sh[2][threadIdx.x + s]);
            sh[3][threadIdx.x] = fmax(sh[3][threadIdx.x], sh[3][threadIdx.x + s]);
            sh[4][threadIdx.x] += sh[4][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x < 5) out[blockIdx.x * 5 + threadIdx.x] = sh[threadIdx.x][0];
}

// Source sort key: pb * ncells + cell, or key_space for a dropped source.
template <typename T>
__global__ __launch_bounds__(kBlk) void k_src_keys(int64_t n, const T *__restrict__ x, const T *__restrict__ y,
                                                   const T *__restrict__ z, const int32_t *__restrict__ patch,
                                                   int32_t n_pb, Grid g, uint32_t *__restrict__ keys,
                                                   uint32_t *__restrict__ idx, uint32_t *__restrict__ counts,
                                                   int32_t *__restrict__ err)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t drop_key = (uint32_t)(g.ncells * n_pb);
    uint32_t key;
    int32_t pb = patch ? patch[i] : 0;
    const bool bad = pb < 0 || pb >= n_pb;
    wave_flag(err, 1, bad);
    if (bad) pb = 0;
    if (z && z[i] < (T)0) {
        key = drop_key;
    } else {
        const int32_t cx = cell_of((double)x[i], g.x0, g.inv, g.nx);
        const int32_t cy = cell_of((double)y[i], g.y0, g.inv, g.ny);
        key = (uint32_t)((int64_t)pb * g.ncells + (int64_t)cy * g.nx + cx);
    }
    keys[i] = key;
    idx[i] = (uint32_t)i;
    wave_count(counts, key, true);
}

// Number of segments of each key: runs longer than kSegMax are split.
__global__ void k_nsegs(int64_t key_space, const uint32_t *__restrict__ counts, uint32_t *__restrict__ nseg)
{
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k <= key_space; k += (int64_t)gridDim.x * blockDim.x)
        nseg[k] = (k < key_space) ? (counts[k] + kSegMax - 1) / kSegMax : 0u;
}

template <typename T>
__global__ __launch_bounds__(kBlk) void k_src_gather(int64_t n, int kind, const uint32_t *__restrict__ order,
                                                     const T *__restrict__ x, const T *__restrict__ y,
                                                     const T *__restrict__ z, const T *__restrict__ s1,
                                                     const T *__restrict__ s2, const T *__restrict__ w,
                                                     const T *__restrict__ m, SrcRec<T> *__restrict__ rec,
                                                     T *__restrict__ zs, int32_t *__restrict__ flags, T zsign)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = order[i];
    wave_flag(flags, 4, kind == SHEAR && m && m[j] != (T)0);
    // a kept source that its values now mask (z < 0, src/tpcf.f90:306-309):
    // only possible when refresh_values re-reads changed values
    wave_flag(flags, 8, z && z[j] < (T)0);
    SrcRec<T> r;
    r.x = x[j];
    r.y = y[j];
    r.w = w[j];
    r.wk = (kind == SHEAR && m) ? w[j] * ((T)1 + m[j]) : w[j];
    r.s1w = s1[j] * r.w;
    r.s2w = s2[j] * r.w;
    r.s1w2 = (T)2 * r.s1w;
    r.s2w2 = (T)2 * r.s2w;
    rec[i] = r;
    if (zs) zs[i] = zsign * z[j];
}

template <typename T>
__global__ __launch_bounds__(kBlk) void k_segs(int64_t key_space, Grid g, const uint32_t *__restrict__ counts,
                                               const uint32_t *__restrict__ cstart, const uint32_t *__restrict__ seg_first,
                                               const SrcRec<T> *__restrict__ rec, const T *__restrict__ zs,
                                               Seg<T> *__restrict__ segs, int4 *__restrict__ pbox)
{
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    // every lane stays to the end (the wave reduction below reads them all)
    const uint32_t total = k < key_space ? counts[k] : 0u;
    const bool live = total > 0;
    const uint32_t pb = live ? (uint32_t)(k / g.ncells) : 0u;
    for (uint32_t part = 0; part * kSegMax < total; ++part) {
        Seg<T> s;
        s.start = cstart[k] + part * kSegMax;
        s.count = min(kSegMax, total - part * kSegMax);
        s.pb = pb;
        s.pad = 0;
        T x0 = rec[s.start].x, x1 = x0, y0 = rec[s.start].y, y1 = y0;
        T zmin = zs ? zs[s.start] : (T)0;
        double W = 0, K = 0, S1 = 0, S2 = 0;
        for (uint32_t q = s.start; q < s.start + s.count; ++q) {
            const SrcRec<T> r = rec[q];
            x0 = r.x < x0 ? r.x : x0; x1 = r.x > x1 ? r.x : x1;
            y0 = r.y < y0 ? r.y : y0; y1 = r.y > y1 ? r.y : y1;
            if (zs) zmin = fmin(zmin, zs[q]);
            W += (double)r.w; K += (double)r.wk; S1 += (double)r.s1w; S2 += (double)r.s2w;
        }
        s.x0 = x0; s.x1 = x1; s.y0 = y0; s.y1 = y1;
        s.zmin = zmin; s.zpad = (T)0;
        s.W = W; s.K = K; s.S1 = S1; s.S2 = S2;
        segs[seg_first[k] + part] = s;
    }
    const int64_t cell = live ? k % g.ncells : 0;
    const int32_t cx = (int32_t)(cell % g.nx), cy = (int32_t)(cell / g.nx);
    // the patch's cell box: one set of atomics per wave when its live keys
    // share a patch (always with one source patch), else one per key
    const uint64_t act = __ballot(live);
    if (!act) return;
    const int lead = __ffsll((unsigned long long)act) - 1;
    const uint32_t pb0 = __shfl(pb, lead);
    if (__ballot(live && pb != pb0) == 0) {
        int32_t x0 = live ? cx : INT32_MAX, x1 = live ? cx : INT32_MIN;
        int32_t y0 = live ? cy : INT32_MAX, y1 = live ? cy : INT32_MIN;
        for (int off = 32; off > 0; off >>= 1) {
            x0 = min(x0, __shfl_xor(x0, off));
            x1 = max(x1, __shfl_xor(x1, off));
            y0 = min(y0, __shfl_xor(y0, off));
            y1 = max(y1, __shfl_xor(y1, off));
        }
        if ((int)__lane_id() == lead) {
            int4 *b = &pbox[pb0];
            atomicMin(&b->x, x0);
            atomicMax(&b->y, x1);
            atomicMin(&b->z, y0);
            atomicMax(&b->w, y1);
        }
    } else if (live) {
        int4 *b = &pbox[pb];
        atomicMin(&b->x, cx);
        atomicMax(&b->y, cx);
        atomicMin(&b->z, cy);
        atomicMax(&b->w, cy);
    }
}

// Lens sort key: (pa << 32) | Hilbert index of the 16-bit quantised position,
// or (n_pa << 32) for a dropped lens.
template <typename T>
__global__ __launch_bounds__(kBlk) void k_lens_keys(int64_t n, const T *__restrict__ x, const T *__restrict__ y,
                                                    const T *__restrict__ z, const T *__restrict__ w,
                                                    const int32_t *__restrict__ patch, int32_t n_pa, double lx0,
                                                    double ly0, double qs, uint64_t *__restrict__ keys,
                                                    uint32_t *__restrict__ idx, uint32_t *__restrict__ pcount,
                                                    int32_t *__restrict__ err)
{
    // per-patch counts: a block histogram in LDS when the patches fit (labels
    // in input order are scattered, so per-lane global atomics would meet on
    // n_pa addresses), else wave-aggregated global atomics
    constexpr int kHist = 1024;
    __shared__ uint32_t hist[kHist];
    const bool use_lds = n_pa <= kHist;
    if (use_lds) {
        for (int b = threadIdx.x; b < n_pa; b += blockDim.x) hist[b] = 0u;
        __syncthreads();
    }
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const bool live = i < n;
    int32_t pa = (live && patch) ? patch[i] : 0;
    const bool bad = live && (pa < 0 || pa >= n_pa);
    wave_flag(err, 2, bad);
    if (bad) pa = 0;
    const bool drop = !live || (w[i] == (T)0) || (z && z[i] < (T)0);
    if (live) {
        uint64_t key;
        if (drop) {
            key = (uint64_t)n_pa << 32;
        } else {
            double fx = ((double)x[i] - lx0) * qs, fy = ((double)y[i] - ly0) * qs;
            uint32_t qx = (uint32_t)fmin(fmax(fx, 0.0), 65535.0);
            uint32_t qy = (uint32_t)fmin(fmax(fy, 0.0), 65535.0);
            key = ((uint64_t)pa << 32) | (uint64_t)hilbert16(qx, qy);
        }
        keys[i] = key;
        idx[i] = (uint32_t)i;
    }
    if (use_lds) {
        if (!drop) atomicAdd(&hist[pa], 1u);
        __syncthreads();
        for (int b = threadIdx.x; b < n_pa; b += blockDim.x)
            if (hist[b]) atomicAdd(&pcount[b], hist[b]);
    } else {
        wave_count(pcount, (uint32_t)pa, !drop);
    }
}

template <typename T>
__global__ __launch_bounds__(kBlk) void k_lens_gather(int64_t n, const uint32_t *__restrict__ order,
                                                      const T *__restrict__ x, const T *__restrict__ y,
                                                      const T *__restrict__ z, const T *__restrict__ val,
                                                      const T *__restrict__ w, T *__restrict__ lx, T *__restrict__ ly,
                                                      T *__restrict__ lz, T *__restrict__ la, T *__restrict__ lw,
                                                      T zsign, int32_t *__restrict__ flags)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = order[i];
    // a kept lens that its values now mask (w == 0, src/tpcf.f90:285 / :474;
    // z < 0 for shear): only possible when refresh_values re-reads changed values
    wave_flag(flags, 16, w[j] == (T)0 || (z && z[j] < (T)0));
    lx[i] = x[j];
    ly[i] = y[j];
    lz[i] = z ? zsign * z[j] : (T)0;
    la[i] = val[j] * w[j];
    lw[i] = w[j];
}

// Greedy cut of one block of sorted lenses (same patch, <= kLensBlock) into
// chunks of at most 64 lenses whose bounding box side stays <= dmax (a jump of
// the Hilbert order out of and back into a patch starts a new chunk): a chunk
// starts at the first lens not yet taken and takes the lenses that follow
// while its box stays within dmax.  One wave per block.  The wave loads the
// next kChunk lenses, one per lane, and forms prefix minima and maxima of x
// and y across the lanes (exact, so these are the boxes a lens-by-lens walk
// grows); the first lane whose prefix box is too wide, or that lies past the
// block, ends the chunk (ballot).  !WRITE counts each block's chunks; WRITE
// stores their lens range, patch and box, and k_chunk_sums then adds the
// value-dependent fields.
template <typename T, bool WRITE>
__global__ __launch_bounds__(kBlk) void k_chunk_blocks(int64_t n_blocks, const uint4 *__restrict__ blocks, T dmax,
                                                       const T *__restrict__ lx, const T *__restrict__ ly,
                                                       uint32_t *__restrict__ nch, const uint32_t *__restrict__ ch_off,
                                                       Chunk<T> *__restrict__ chunks)
{
    static_assert(kChunk == kWave, "one lens per lane");
    const int64_t bi = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / kWave;   // wave-uniform
    const int lane = (int)(threadIdx.x % kWave);
    if (bi >= n_blocks) return;
    const uint4 blk = blocks[bi];   // start, count, pa
    const uint32_t end = blk.x + blk.y;
    uint32_t s = blk.x, c = 0;
    while (s < end) {
        const uint32_t i = s + (uint32_t)lane;
        const bool in = i < end;
        T x0 = lx[in ? i : s], y0 = ly[in ? i : s];
        T x1 = x0, y1 = y0;
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            const T ax0 = __shfl_up(x0, o), ax1 = __shfl_up(x1, o);
            const T ay0 = __shfl_up(y0, o), ay1 = __shfl_up(y1, o);
            if (lane >= o) {
                x0 = ax0 < x0 ? ax0 : x0;
                x1 = ax1 > x1 ? ax1 : x1;
                y0 = ay0 < y0 ? ay0 : y0;
                y1 = ay1 > y1 ? ay1 : y1;
            }
        }
        const uint64_t bad = __builtin_amdgcn_ballot_w64(!in || x1 - x0 > dmax || y1 - y0 > dmax);
        // lane 0 is in the block and its box is a point, so n >= 1
        const int n = bad ? __builtin_ctzll(bad) : kWave;
        if (WRITE) {
            const T bx0 = __shfl(x0, n - 1), bx1 = __shfl(x1, n - 1);
            const T by0 = __shfl(y0, n - 1), by1 = __shfl(y1, n - 1);
            if (lane == 0) {
                Chunk<T> ch;
                ch.start = s;
                ch.count = (uint32_t)n;
                ch.pa = blk.z;
                ch.pad = 0;
                ch.bx0 = bx0; ch.bx1 = bx1; ch.by0 = by0; ch.by1 = by1;
                ch.zmax = ch.zpad = (T)0;
                ch.wsum = ch.asum = 0.0;
                chunks[ch_off[bi] + c] = ch;
            }
        }
        s += (uint32_t)n;
        ++c;
    }
    if (!WRITE && lane == 0) nch[bi] = c;
}

// Value-dependent fields of chunks (z maximum, sums of w and val w), one
// thread per chunk: after k_chunk_blocks<T, true>, and again when new values
// arrive on an unchanged structure.
template <typename T>
__global__ __launch_bounds__(kBlk) void k_chunk_sums(int64_t n_chunks, const T *__restrict__ lz,
                                                     const T *__restrict__ lw, const T *__restrict__ la,
                                                     Chunk<T> *__restrict__ chunks)
{
    const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (c >= n_chunks) return;
    const uint32_t s = chunks[c].start, e = s + chunks[c].count;
    T zmax = lz[s];
    double ws = 0, as = 0;
    for (uint32_t q = s; q < e; ++q) {
        zmax = fmax(zmax, lz[q]);
        ws += (double)lw[q];
        as += (double)la[q];
    }
    chunks[c].zmax = zmax;
    chunks[c].wsum = ws;
    chunks[c].asum = as;
}

// Sources a chunk can meet (the sum over its candidate row ranges) and the
// cell rows it can reach.
template <typename T>
__global__ __launch_bounds__(kBlk) void k_chunk_work(int64_t n_chunks, Grid g, int32_t n_pb, double xm, double mg,
                                                     const Chunk<T> *__restrict__ chunks,
                                                     const int4 *__restrict__ pbox, const uint32_t *__restrict__ cstart,
                                                     float *__restrict__ work, int2 *__restrict__ rows)
{
    const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (c >= n_chunks) return;
    const Chunk<T> ch = chunks[c];
    int32_t ry0, ry1;
    row_span(g, (double)ch.by0, (double)ch.by1, xm, mg, ry0, ry1);
    rows[c] = make_int2(ry0, ry1);
    double total = 0.0;
    for (int32_t pb = 0; pb < n_pb; ++pb) {
        const int4 b = pbox[pb];
        if (b.x > b.y) continue;
        const int32_t cy0 = max(ry0, b.z), cy1 = min(ry1, b.w);
        for (int32_t cy = cy0; cy <= cy1; ++cy) {
            int32_t cx0, cx1;
            if (!col_span(g, (double)ch.bx0, (double)ch.bx1, (double)ch.by0, (double)ch.by1, cy, xm, mg, cx0, cx1))
                continue;
            cx0 = max(cx0, b.x);
            cx1 = min(cx1, b.y);
            if (cx0 > cx1) continue;
            const int64_t key0 = (int64_t)pb * g.ncells + (int64_t)cy * g.nx + cx0;
            total += (double)(cstart[key0 + (cx1 - cx0) + 1] - cstart[key0]);
        }
    }
    work[c] = (float)total;
}

inline unsigned blocks_for(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + kBlk - 1) / kBlk); }

// The pair kernel's work items.  A chunk's reachable rows [ry0, ry1] are cut
// at multiples of a band height of split * x_max (whole cell rows, at least
// one), one item per (band, chunk) with work estimated as the chunk's sources
// times the item's share of its rows.  The items are ordered band by band
// (chunks in their sorted order within a band) and the sequence is cut into
// kSlots consecutive ranges of equal estimated work, one per XCD: the pair
// kernel deals range s to XCD s in order, so the waves running on an XCD
// together cover one band of source rows for neighbouring chunks, and the
// rows they read stay in that XCD's L2.  Each range is padded with kNoChunk
// items to a common length (a multiple of the waves per workgroup).  Chunks
// that meet no source get no item.
std::vector<Item> band_items(const std::vector<float> &work, const std::vector<int2> &rows, const Grid &g,
                             double x_max, double split)
{
    const int64_t nch = (int64_t)work.size();
    const int band = (int)std::max(1.0, std::min(65535.0, std::floor(split * x_max / g.cell + 0.5)));
    const int n_bands = (int)((g.ny + band - 1) / band);
    // counting sort of the items by band
    std::vector<int64_t> first(n_bands + 1, 0);
    for (int64_t c = 0; c < nch; ++c) {
        if (!(work[c] > 0.0f)) continue;
        for (int b = rows[c].x / band; b <= rows[c].y / band; ++b) ++first[b + 1];
    }
    for (int b = 0; b < n_bands; ++b) first[b + 1] += first[b];
    const int64_t n = first[n_bands];
    std::vector<Item> seq(n);
    std::vector<float> wseq(n);
    std::vector<int64_t> at(first.begin(), first.end() - 1);
    for (int64_t c = 0; c < nch; ++c) {
        if (!(work[c] > 0.0f)) continue;
        const int r0 = rows[c].x, r1 = rows[c].y;
        for (int b = r0 / band; b <= r1 / band; ++b) {
            const int lo = std::max(r0, b * band), hi = std::min(r1, b * band + band - 1);
            const int64_t k = at[b]++;
            seq[k] = Item{(uint32_t)c, (uint16_t)lo, (uint16_t)hi};
            wseq[k] = work[c] * (float)(hi - lo + 1) / (float)(r1 - r0 + 1);
        }
    }
    if (n == 0) return {};
    // kSlots consecutive ranges of equal estimated work
    double tot = 0.0;
    for (float v : wseq) tot += v;
    std::vector<int64_t> cut(kSlots + 1, n);
    cut[0] = 0;
    double acc = 0.0;
    int s = 1;
    for (int64_t k = 0; k < n && s < kSlots; ++k) {
        acc += wseq[k];
        while (s < kSlots && acc >= tot * s / kSlots) cut[s++] = k + 1;
    }
    int64_t len = 0;
    for (int q = 0; q < kSlots; ++q) len = std::max(len, cut[q + 1] - cut[q]);
    len = (len + kWavesPerBlock - 1) / kWavesPerBlock * kWavesPerBlock;
    std::vector<Item> out((size_t)(len * kSlots), Item{kNoChunk, 0, 0});
    for (int q = 0; q < kSlots; ++q)
        std::copy(seq.begin() + cut[q], seq.begin() + cut[q + 1], out.begin() + q * len);
    return out;
}

inline int bitlen(uint64_t v)
{
    int b = 0;
    while (v) { ++b; v >>= 1; }
    return b;
}

}  // namespace

template <typename T>
int prepare(tpcf_ctx *ctx, const CallArgs<T> &a, Prepared<T> &p)
{
    hipStream_t s = ctx->stream;
    const bool shear = (a.kind == SHEAR);
    const double x_max = (double)(a.kbins ? a.kbins : a.bins)->x_max;
    p.sph_lp = p.sph_lf = nullptr;
    p.sph_segp = p.sph_chp = nullptr;
    p.sph_rev = false;

    // ---- bounding boxes and finiteness of both catalogues ----
    const unsigned nbs = (unsigned)std::min<int64_t>(1024, (a.n_b + kBlk - 1) / kBlk);
    const unsigned nbl = (unsigned)std::min<int64_t>(1024, (a.n_a + kBlk - 1) / kBlk);
    double *partial;
    TPCF_TRY(ensure(ctx, tpcf_ctx::B_PARTIAL, sizeof(double) * 5 * (nbs + nbl), (void **)&partial));
    hipLaunchKernelGGL(k_bbox<T>, dim3(nbs), dim3(kBlk), 0, s, a.n_b, a.xb, a.yb, partial);
    hipLaunchKernelGGL(k_bbox<T>, dim3(nbl), dim3(kBlk), 0, s, a.n_a, a.xa, a.ya, partial + 5 * nbs);
    std::vector<double> hp(5 * (nbs + nbl));
    TPCF_HIP(ctx, hipMemcpyAsync(hp.data(), partial, sizeof(double) * hp.size(), hipMemcpyDeviceToHost, s));
    TPCF_HIP(ctx, hipStreamSynchronize(s));
    double sb[5] = {INFINITY, -INFINITY, INFINITY, -INFINITY, 0}, lb[5] = {INFINITY, -INFINITY, INFINITY, -INFINITY, 0};
    for (unsigned b = 0; b < nbs + nbl; ++b) {
        double *d = (b < nbs) ? sb : lb;
        const double *q = &hp[5 * b];
        d[0] = std::min(d[0], q[0]); d[1] = std::max(d[1], q[1]);
        d[2] = std::min(d[2], q[2]); d[3] = std::max(d[3], q[3]);
        d[4] += q[4];
    }
    if (sb[4] > 0 || lb[4] > 0) return fail(ctx, TPCF_ERR_NONFINITE, "non-finite coordinate in the input catalogues");

    // ---- source cell grid ----
    const double ext_x = sb[1] - sb[0], ext_y = sb[3] - sb[2];
    double cell = ctx->opt_cell;
    if (!(cell > 0.0)) {
        const double floor_ext = x_max / 64.0;
        const double area = std::max(ext_x, floor_ext) * std::max(ext_y, floor_ext);
        cell = std::sqrt(cell_target<T>(ctx) * area / (double)std::max<int64_t>(a.n_b, 1));
        cell = std::min(std::max(cell, x_max / 512.0), 4.0 * x_max);
    }
    Grid g;
    for (int attempt = 0;; ++attempt) {
        cell = std::max(cell, std::max(ext_x, ext_y) / 16384.0);
        g.x0 = sb[0];
        g.y0 = sb[2];
        g.cell = cell;
        g.inv = 1.0 / cell;
        g.nx = (int32_t)std::floor(ext_x * g.inv) + 1;
        g.ny = (int32_t)std::floor(ext_y * g.inv) + 1;
        g.ncells = (int64_t)g.nx * g.ny;
        if (g.ncells * a.n_pb < ((int64_t)1 << 31)) break;
        if (attempt > 40) return fail(ctx, TPCF_ERR_GRID, "cell grid too large");
        cell *= 1.25;
    }
    const int64_t key_space = g.ncells * a.n_pb;
    p.grid = g;
    p.key_space = key_space;

    // ---- sources: keys, histogram, sort ----
    const int64_t nb = a.n_b;
    uint32_t *skey, *skey2, *sidx, *sidx2, *counts, *cstart, *seg_first;
    int32_t *err;
    TPCF_TRY(ensure(ctx, tpcf_ctx::B_SKEY, sizeof(uint32_t) * nb, (void **)&skey));
    TPCF_TRY(ensure(ctx, tpcf_ctx::B_SKEY2, sizeof(uint32_t) * nb, (void **)&skey2));
    TPCF_TRY(ensure(ctx, tpcf_ctx::B_SIDX, sizeof(uint32_t) * nb, (void **)&sidx));
    TPCF_TRY(ensure(ctx, tpcf_ctx::B_SIDX2, sizeof(uint32_t) * nb, (void **)&sidx2));
    TPCF_TRY(ensure(ctx, tpcf_ctx::B_COUNTS, sizeof(uint32_t) * (key_space + 2), (void **)&counts));
    TPCF_TRY(ensure(ctx, tpcf_ctx::B_CSTART, sizeof(uint32_t) * (key_space + 2), (void **)&cstart));
    TPCF_TRY(ensure(ctx, tpcf_ctx::B_SEGFIRST, sizeof(uint32_t) * (key_space + 2), (void **)&seg_first));
    TPCF_TRY(ensure(ctx, tpcf_ctx::B_ERR, sizeof(int32_t) * 4, (void **)&err));
    launch_fill<uint32_t>(s, counts, 0u, key_space + 1);
    launch_fill<int32_t>(s, err, 0, 4);
    hipLaunchKernelGGL(k_src_keys<T>, dim3(blocks_for(nb)), dim3(kBlk), 0, s, nb, a.xb, a.yb, shear ? a.zb : nullptr,
                       a.patch_b, a.n_pb, g, skey, sidx, counts, err);

    size_t tmp_sort = 0, tmp_scan = 0, tmp_lsort = 0;
    const int sbits = std::max(1, bitlen((uint64_t)key_space));
    TPCF_HIP(ctx, rocprim::radix_sort_pairs(nullptr, tmp_sort, skey, skey2, sidx, sidx2, (size_t)nb, 0, sbits, s));
    TPCF_HIP(ctx, rocprim::exclusive_scan(nullptr, tmp_scan, counts, cstart, 0u, (size_t)key_space + 1,
                                          rocprim::plus<uint32_t>(), s));
    const int lbits = 32 + std::max(1, bitlen((uint64_t)a.n_pa));
    {
        uint64_t *k64 = nullptr;
        TPCF_HIP(ctx, rocprim::radix_sort_pairs(nullptr, tmp_lsort, k64, k64, sidx, sidx2, (size_t)a.n_a, 0, lbits, s));
    }
    void *tmp;
    TPCF_TRY(ensure(ctx, tpcf_ctx::B_TMP, std::max(std::max(tmp_sort, tmp_scan), tmp_lsort) + 256, &tmp));
    size_t tb = tmp_sort;
    TPCF_HIP(ctx, rocprim::radix_sort_pairs(tmp, tb, skey, skey2, sidx, sidx2, (size_t)nb, 0, sbits, s));
    tb = tmp_scan;
    TPCF_HIP(ctx, rocprim::exclusive_scan(tmp, tb, counts, cstart, 0u, (size_t)key_space + 1,
                                          rocprim::plus<uint32_t>(), s));
    // segments per key, scanned into the first segment index of each key
    uint32_t *flags;
    TPCF_TRY(ensure(ctx, tpcf_ctx::B_OUT, sizeof(uint32_t) * (key_space + 2), (void **)&flags));
    hipLaunchKernelGGL(k_nsegs, dim3((unsigned)std::min<int64_t>(8192, blocks_for(key_space + 1))), dim3(kBlk), 0, s,
                       key_space, counts, flags);
    tb = tmp_scan;
    TPCF_HIP(ctx, rocprim::exclusive_scan(tmp, tb, flags, seg_first, 0u, (size_t)key_space + 1,
                                          rocprim::plus<uint32_t>(), s));
    uint32_t hcnt[2];
    int32_t herr = 0;
    TPCF_HIP(ctx, hipMemcpyAsync(&hcnt[0], cstart + key_space, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    TPCF_HIP(ctx, hipMemcpyAsync(&hcnt[1], seg_first + key_space, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    TPCF_HIP(ctx, hipMemcpyAsync(&herr, err, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    TPCF_HIP(ctx, hipStreamSynchronize(s));
    if (herr & 1) return fail(ctx, TPCF_ERR_PATCH, "source patch label outside [0, n_patch_B)");
    p.n_src = hcnt[0];
    p.n_segs = hcnt[1];

    SrcRec<T> *srec;
    T *sz = nullptr;
    Seg<T> *segs;
    int4 *pbox;
    TPCF_TRY(ensure(ctx, tpcf_ctx::B_SREC, sizeof(SrcRec<T>) * (p.n_src + kRecPad), (void **)&srec));
    launch_fill<T>(s, reinterpret_cast<T *>(srec + p.n_src), (T)0, 8 * kRecPad);
    if (shear) TPCF_TRY(ensure(ctx, tpcf_ctx::B_SZ, sizeof(T) * std::max<int64_t>(p.n_src, 1), (void **)&sz));
    TPCF_TRY(ensure(ctx, tpcf_ctx::B_SEGS, sizeof(Seg<T>) * std::max<int64_t>(p.n_segs, 1), (void **)&segs));
    TPCF_TRY(ensure(ctx, tpcf_ctx::B_PBOX, sizeof(int4) * a.n_pb, (void **)&pbox));
    if (p.n_src > 0)
        hipLaunchKernelGGL(k_src_gather<T>, dim3(blocks_for(p.n_src)), dim3(kBlk), 0, s, p.n_src, a.kind, sidx2, a.xb,
                           a.yb, a.zb, a.s1b, a.s2b, a.wb, a.mb, srec, sz, err, (T)a.zsign);
    launch_fill<int4>(s, pbox, make_int4(INT_MAX, -1, INT_MAX, -1), a.n_pb);
    hipLaunchKernelGGL(k_segs<T>, dim3(blocks_for(key_space)), dim3(kBlk), 0, s, key_space, g, counts, cstart,
                       seg_first, srec, sz, segs, pbox);
    p.srec = srec;
    p.sz = sz;
    p.segs = segs;
    p.seg_first = seg_first;
    p.pbox = pbox;

    // ---- lenses: keys, sort, chunks ----
    const int64_t na = a.n_a;
    uint64_t *lkey, *lkey2;
    uint32_t *lidx, *lidx2, *pcount;
    TPCF_TRY(ensure(ctx, tpcf_ctx::B_LKEY, sizeof(uint64_t) * na, (void **)&lkey));
    TPCF_TRY(ensure(ctx, tpcf_ctx::B_LKEY2, sizeof(uint64_t) * na, (void **)&lkey2));
    TPCF_TRY(ensure(ctx, tpcf_ctx::B_LIDX, sizeof(uint32_t) * na, (void **)&lidx));
    TPCF_TRY(ensure(ctx, tpcf_ctx::B_LIDX2, sizeof(uint32_t) * na, (void **)&lidx2));
    TPCF_TRY(ensure(ctx, tpcf_ctx::B_PCOUNT, sizeof(uint32_t) * (a.n_pa + 1), (void **)&pcount));
    launch_fill<uint32_t>(s, pcount, 0u, a.n_pa + 1);
    const double lext = std::max(std::max(lb[1] - lb[0], lb[3] - lb[2]), 1e-30);
    hipLaunchKernelGGL(k_lens_keys<T>, dim3(blocks_for(na)), dim3(kBlk), 0, s, na, a.xa, a.ya, shear ? a.za : nullptr,
                       a.wa, a.patch_a, a.n_pa, lb[0], lb[2], 65535.0 / lext, lkey, lidx, pcount, err);
    tb = tmp_lsort;
    TPCF_HIP(ctx, rocprim::radix_sort_pairs(tmp, tb, lkey, lkey2, lidx, lidx2, (size_t)na, 0, lbits, s));
    std::vector<uint32_t> hpc(a.n_pa);
    TPCF_HIP(ctx, hipMemcpyAsync(hpc.data(), pcount, sizeof(uint32_t) * a.n_pa, hipMemcpyDeviceToHost, s));
    TPCF_HIP(ctx, hipMemcpyAsync(&herr, err, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    TPCF_HIP(ctx, hipStreamSynchronize(s));
    if (herr & 2) return fail(ctx, TPCF_ERR_PATCH, "lens patch label outside [0, n_patch_A)");
    // blocks of <= kLensBlock sorted lenses within one patch
    constexpr uint32_t kLensBlock = 4096;
    std::vector<uint4> hblk;
    int64_t nl = 0;
    for (int32_t q = 0; q < a.n_pa; ++q) {
        for (uint32_t o = 0; o < hpc[q]; o += kLensBlock)
            hblk.push_back(make_uint4((uint32_t)nl + o, std::min(kLensBlock, hpc[q] - o), (uint32_t)q, 0u));
        nl += hpc[q];
    }
    p.n_lens = nl;
    const int64_t nblk = (int64_t)hblk.size();
    uint4 *dblk;
    uint32_t *dnch;
    TPCF_TRY(ensure(ctx, tpcf_ctx::B_BLK, sizeof(uint4) * std::max<int64_t>(nblk, 1), (void **)&dblk));
    TPCF_TRY(ensure(ctx, tpcf_ctx::B_BLKCH, sizeof(uint32_t) * 2 * std::max<int64_t>(nblk, 1), (void **)&dnch));
    if (nblk > 0)
        TPCF_HIP(ctx, hipMemcpyAsync(dblk, hblk.data(), sizeof(uint4) * nblk, hipMemcpyHostToDevice, s));

    T *lx, *ly, *lz, *la, *lw;
    Chunk<T> *chunks;
    const int64_t nl1 = std::max<int64_t>(nl, 1);
    TPCF_TRY(ensure(ctx, tpcf_ctx::B_LX, sizeof(T) * nl1, (void **)&lx));
    TPCF_TRY(ensure(ctx, tpcf_ctx::B_LY, sizeof(T) * nl1, (void **)&ly));
    TPCF_TRY(ensure(ctx, tpcf_ctx::B_LZ, sizeof(T) * nl1, (void **)&lz));
    TPCF_TRY(ensure(ctx, tpcf_ctx::B_LA, sizeof(T) * nl1, (void **)&la));
    TPCF_TRY(ensure(ctx, tpcf_ctx::B_LW, sizeof(T) * nl1, (void **)&lw));
    // largest chunk box side: three times the side holding 64 lenses at the mean density
    const double lens_area = std::max((lb[1] - lb[0]) * (lb[3] - lb[2]), 1e-30);
    const T dmax = (T)(3.0 * std::sqrt((double)kChunk * lens_area / (double)std::max<int64_t>(nl, 1)));
    int64_t nch = 0;
    std::vector<uint32_t> hnch(nblk), hchoff(nblk);
    if (nl > 0) {
        hipLaunchKernelGGL(k_lens_gather<T>, dim3(blocks_for(nl)), dim3(kBlk), 0, s, nl, lidx2, a.xa, a.ya,
                           shear ? a.za : nullptr, a.vala, a.wa, lx, ly, lz, la, lw, (T)a.zsign, err);
        hipLaunchKernelGGL((k_chunk_blocks<T, false>), dim3(blocks_for(nblk * kWave)), dim3(kBlk), 0, s, nblk, dblk,
                           dmax, lx, ly, dnch, nullptr, nullptr);
        TPCF_HIP(ctx, hipMemcpyAsync(hnch.data(), dnch, sizeof(uint32_t) * nblk, hipMemcpyDeviceToHost, s));
        TPCF_HIP(ctx, hipMemcpyAsync(&herr, err, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        TPCF_HIP(ctx, hipStreamSynchronize(s));
        for (int64_t bi = 0; bi < nblk; ++bi) {
            hchoff[bi] = (uint32_t)nch;
            nch += hnch[bi];
        }
        TPCF_HIP(ctx, hipMemcpyAsync(dnch + nblk, hchoff.data(), sizeof(uint32_t) * nblk, hipMemcpyHostToDevice, s));
    }
    p.n_chunks = nch;
    p.m_zero = !(herr & 4);
    TPCF_TRY(ensure(ctx, tpcf_ctx::B_CHUNKS, sizeof(Chunk<T>) * std::max<int64_t>(nch, 1), (void **)&chunks));
    p.margin = 1e-9 * (std::fabs(g.x0) + std::fabs(g.y0) + g.cell * (g.nx + g.ny) + x_max) + 1e-12;
    Item *items;
    if (nch > 0) {
        hipLaunchKernelGGL((k_chunk_blocks<T, true>), dim3(blocks_for(nblk * kWave)), dim3(kBlk), 0, s, nblk, dblk,
                           dmax, lx, ly, nullptr, dnch + nblk, chunks);
        hipLaunchKernelGGL(k_chunk_sums<T>, dim3(blocks_for(nch)), dim3(kBlk), 0, s, nch, lz, lw, la, chunks);
        // each chunk's reachable source rows and the sources in them
        float *dwork;
        TPCF_TRY(ensure(ctx, tpcf_ctx::B_WORK, (sizeof(float) + sizeof(int2)) * nch + 16, (void **)&dwork));
        int2 *drows = reinterpret_cast<int2 *>(reinterpret_cast<char *>(dwork) + ((sizeof(float) * nch + 15) & ~(size_t)15));
        hipLaunchKernelGGL(k_chunk_work<T>, dim3(blocks_for(nch)), dim3(kBlk), 0, s, nch, g, a.n_pb, x_max, p.margin,
                           chunks, pbox, cstart, dwork, drows);
        std::vector<float> hw(nch);
        std::vector<int2> hrows(nch);
        TPCF_HIP(ctx, hipMemcpyAsync(hw.data(), dwork, sizeof(float) * nch, hipMemcpyDeviceToHost, s));
        TPCF_HIP(ctx, hipMemcpyAsync(hrows.data(), drows, sizeof(int2) * nch, hipMemcpyDeviceToHost, s));
        TPCF_HIP(ctx, hipStreamSynchronize(s));
        std::vector<Item> hit = band_items(hw, hrows, g, x_max, ctx->opt_split);
        p.n_items = (int64_t)hit.size();
        TPCF_TRY(ensure(ctx, tpcf_ctx::B_ITEMS, sizeof(Item) * std::max<size_t>(hit.size(), 1), (void **)&items));
        if (!hit.empty())
            TPCF_HIP(ctx, hipMemcpyAsync(items, hit.data(), sizeof(Item) * hit.size(), hipMemcpyHostToDevice, s));
        TPCF_HIP(ctx, hipStreamSynchronize(s));
        p.items = items;
    } else {
        p.n_items = 0;
        p.items = nullptr;
    }
    p.cstart = cstart;
    p.sidx = sidx2;
    p.counts = counts;
    p.err = err;
    TPCF_HIP(ctx, hipGetLastError());
    p.chunks = chunks;
    p.lx = lx; p.ly = ly; p.lz = lz; p.la = la; p.lw = lw;
    p.lidx = lidx2;
    return TPCF_OK;
}

template <typename T>
int refresh_values(tpcf_ctx *ctx, const CallArgs<T> &a, const Prepared<T> &p, hipStream_t s)
{
    const bool shear = (a.kind == SHEAR);
    launch_fill<int32_t>(s, p.err, 0, 4);
    if (p.n_src > 0)
        hipLaunchKernelGGL(k_src_gather<T>, dim3(blocks_for(p.n_src)), dim3(kBlk), 0, s, p.n_src, a.kind, p.sidx, a.xb,
                           a.yb, a.zb, a.s1b, a.s2b, a.wb, a.mb, const_cast<SrcRec<T> *>(p.srec),
                           const_cast<T *>(p.sz), p.err, (T)a.zsign);
    launch_fill<int4>(s, const_cast<int4 *>(p.pbox), make_int4(INT_MAX, -1, INT_MAX, -1), a.n_pb);
    hipLaunchKernelGGL(k_segs<T>, dim3(blocks_for(p.key_space)), dim3(kBlk), 0, s, p.key_space, p.grid, p.counts,
                       p.cstart, p.seg_first, p.srec, p.sz, const_cast<Seg<T> *>(p.segs), const_cast<int4 *>(p.pbox));
    if (p.n_lens > 0) {
        hipLaunchKernelGGL(k_lens_gather<T>, dim3(blocks_for(p.n_lens)), dim3(kBlk), 0, s, p.n_lens, p.lidx, a.xa,
                           a.ya, shear ? a.za : nullptr, a.vala, a.wa, const_cast<T *>(p.lx), const_cast<T *>(p.ly),
                           const_cast<T *>(p.lz), const_cast<T *>(p.la), const_cast<T *>(p.lw), (T)a.zsign, p.err);
        if (p.n_chunks > 0)
            hipLaunchKernelGGL(k_chunk_sums<T>, dim3(blocks_for(p.n_chunks)), dim3(kBlk), 0, s, p.n_chunks, p.lz,
                               p.lw, p.la, const_cast<Chunk<T> *>(p.chunks));
    }
    TPCF_HIP(ctx, hipGetLastError());
    return TPCF_OK;
}

template int refresh_values<float>(tpcf_ctx *, const CallArgs<float> &, const Prepared<float> &, hipStream_t);
template int refresh_values<double>(tpcf_ctx *, const CallArgs<double> &, const Prepared<double> &, hipStream_t);

template int prepare<float>(tpcf_ctx *, const CallArgs<float> &, Prepared<float> &);
template int prepare<double>(tpcf_ctx *, const CallArgs<double> &, Prepared<double> &);

// fill helpers used by the other translation units
void fill_u64(hipStream_t s, unsigned long long *p, int64_t n) { launch_fill<unsigned long long>(s, p, 0ull, n); }
void fill_f64(hipStream_t s, double *p, int64_t n) { launch_fill<double>(s, p, 0.0, n); }

}  // namespace tpcf
