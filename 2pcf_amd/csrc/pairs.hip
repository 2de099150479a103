// pairs.hip -- the pair kernel: the inner loops of tangential_shear_tpcf
// (src/tpcf.f90:266-359) and scalar_kappa_tpcf (src/tpcf.f90:456-526).
//
// One wavefront owns one lens chunk (up to 64 lenses, one per lane) and walks
// the source segments within x_max of the chunk's bounding box, source patch
// by source patch, row by row (cell rows of the grid).  Source records are
// wave-uniform, so they are read with scalar loads and every lane pairs its
// lens with the same source.  Each (chunk, segment) pair is classified from
// the two bounding boxes:
//   fast  -- every pair lies in one bin: no per-pair binning; W, K and the
//            pair count come from the segment sums (and G for kappa);
//   edge  -- pairs straddle one bin edge: per-pair FP32 test against the
//            squared edge, FP64 re-test for pairs within the band;
//   multi -- several edges (or a z test): windows of two bins, bounded tests.
// Each segment's pairs are summed per lane in the input precision (at most
// kSegMax terms), then added to per-lane FP64 sums for a window of two
// adjacent bins.  When the window moves these are reduced across the wave
// (DPP) into FP64 per-bin registers (lane b holds bin b), which go to the
// per-patch-pair FP64 accumulators with global atomics when the source patch
// changes.
#include <cmath>

#include "internal.hpp"

namespace tpcf {

void fill_u64(hipStream_t s, unsigned long long *p, int64_t n);
void fill_f64(hipStream_t s, double *p, int64_t n);

template <typename T>
struct PairArgs {
    Grid g;
    int64_t n_chunks;
    int64_t per_xcd;       // blocks per XCD of the remapped grid
    int32_t n_pa, n_pb, n_bin, counters;
    double x_max;          // outer edge (FP64)
    int32_t xmin_zero;     // x_min == 0: pairs at zero separation fall into bin 0
    double margin;         // absolute slack for cell-range bounds
    T cls_lo, cls_hi;      // classification margins on squared box distances
    float band;            // relative ambiguity band on d^2 (FP32 path)
    float e2f[TPCF_MAX_BIN + 1];   // squared edges rounded to FP32 (+inf past n_bin)
    double e2d[TPCF_MAX_BIN + 1];  // squared edges in FP64; e2d[n_bin] = next double above x_max^2
};

template <typename T>
struct Geo;

template <>
struct Geo<float> {
    // separation and 1/d^2 from d^2 > 0 with one reciprocal square root
    static __device__ inline void dinv(float d2, float &d, float &inv)
    {
        const float r = __builtin_amdgcn_rsqf(d2);
        d = d2 * r;
        inv = r * r;
    }
    static __device__ inline float dist(float d2) { return __builtin_amdgcn_sqrtf(d2); }
};

template <>
struct Geo<double> {
    static __device__ inline void dinv(double d2, double &d, double &inv)
    {
        d = sqrt(d2);
        inv = 1.0 / d2;
    }
    static __device__ inline double dist(double d2) { return sqrt(d2); }
};

// d^2 < edge test for the squared edge with index e.  FP32: decided in FP32
// unless d^2 lies within the band around the edge, then re-decided in FP64
// from the FP32 inputs exactly as the FP64 oracle does.  FP64: direct.
template <typename T>
__device__ inline bool below_edge(const PairArgs<T> &A, int e, T d2, T xa, T ya, T xs, T ys)
{
    if constexpr (sizeof(T) == 4) {
        const float ef = A.e2f[e];
        bool p = d2 < ef;
        if (fabsf(d2 - ef) <= A.band * ef) {
            const double dx = (double)xa - (double)xs, dy = (double)ya - (double)ys;
            p = fma(dx, dx, dy * dy) < A.e2d[e];
        }
        return p;
    } else {
        return d2 < A.e2d[e];
    }
}

// Per-wave window state: two adjacent bins (wk, wk+1) of one source patch.
struct Win {
    int pb, wk;
    bool open, used0, used1, edge;
    double aT[2], aG1[2], aG2[2], aW[2], aK[2], aN[2];   // per-lane FP64 sums
    double aZ;   // shear, x_min == 0: in-bin pairs at zero separation (bin 0 = window slot 0)
};

// Per-lane FP64 per-bin sums of the current source patch (lane b holds bin b).
struct Table {
    double T, G1, G2, W, K, N;
};

template <typename T>
struct LaneData {
    T xa, ya, za, wa, aa, vmask;
    bool valid;
    uint32_t li;
    int lane;
};

__device__ __forceinline__ void win_clear(Win &w)
{
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        w.aT[s] = 0; w.aG1[s] = 0; w.aG2[s] = 0; w.aW[s] = 0; w.aK[s] = 0; w.aN[s] = 0;
    }
    w.aZ = 0;
    w.open = false;
    w.used0 = w.used1 = false;
    w.edge = false;
}

// Reduce the window's per-lane sums across the wave into the per-bin table
// (and, with marks, add each lane's sums to its lens' marks).
template <typename T, int KIND, bool MARKS>
__device__ __forceinline__ void win_flush(Win &w, Table &t, const LaneData<T> &L, int n_bin,
                                          double *__restrict__ marks)
{
    if (!w.open) return;
    const double dwa = (double)L.wa, daa = (double)L.aa;
    const double gs = (KIND == SHEAR) ? -1.0 : 1.0;   // G_t -= ... for shear, src/tpcf.f90:348-349
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const int bin = w.wk + s;
        const bool used = s ? w.used1 : w.used0;
        if (!used || bin < 0 || bin >= n_bin) continue;
        const double sT = wave_sum(dwa * w.aT[s]);
        double sG1 = 0, sG2 = 0, sW = 0, sK = 0, sN = 0;
        if (KIND == SHEAR || MARKS || w.edge) {
            sG1 = gs * wave_sum(daa * w.aG1[s]);
            sG2 = gs * wave_sum(daa * w.aG2[s]);
        }
        if (MARKS || w.edge) {
            sW = wave_sum(dwa * w.aW[s]);
            sK = wave_sum(dwa * w.aK[s]);
        }
        if (w.edge) sN = wave_sum((double)L.vmask * w.aN[s]);
        if (KIND == SHEAR && bin == 0 && w.edge && wave_sum((double)L.vmask * w.aZ) > 0.0) {
            // the position angle of a zero-separation pair is 0/0: the reference's
            // G_t and G_x of that bin become NaN (src/tpcf.f90:343-349)
            sG1 = __builtin_nan("");
            sG2 = __builtin_nan("");
        }
        if (L.lane == bin) {
            t.T += sT; t.G1 += sG1; t.G2 += sG2; t.W += sW; t.K += sK; t.N += sN;
        }
        if (MARKS && L.valid) {
            double *mk = marks + ((size_t)L.li * n_bin + bin) * 4;
            mk[0] += gs * daa * w.aG1[s];
            mk[1] += gs * daa * w.aG2[s];
            mk[2] += dwa * w.aK[s];
            mk[3] += dwa * w.aW[s];
        }
    }
    win_clear(w);
}

template <typename T, int KIND, bool MARKS>
__device__ __forceinline__ void win_open(Win &w, Table &t, const LaneData<T> &L, int n_bin, double *marks, int pb,
                                         int k)
{
    if (w.open && (pb != w.pb || k != w.wk)) win_flush<T, KIND, MARKS>(w, t, L, n_bin, marks);
    w.pb = pb;
    w.wk = k;
    w.open = true;
}

// Per-pair pass over a segment for the window bins wk (mask m) and wk+1 (m2).
template <typename T, int KIND, bool BOUNDED, bool ZCHECK>
__device__ __forceinline__ void win_pass(Win &w, const LaneData<T> &L, const PairArgs<T> &A,
                                         const SrcRec<T> *__restrict__ r, const T *__restrict__ zr, uint32_t count,
                                         int n_bin)
{
    const int lo_e = w.wk, mid_e = w.wk + 1, hi_e = w.wk + 2;
    const bool track0 = (KIND == SHEAR) && A.xmin_zero && (w.wk == 0);
    T sT0 = 0, sT1 = 0, sG10 = 0, sG11 = 0, sG20 = 0, sG21 = 0;
    T sW0 = 0, sW1 = 0, sK0 = 0, sK1 = 0, sN0 = 0, sN1 = 0, sZ = 0;
    for (uint32_t j = 0; j < count; ++j) {
        const SrcRec<T> q = r[j];
        const T dx = L.xa - q.x, dy = L.ya - q.y;
        const T dy2 = dy * dy;
        const T d2 = fma(dx, dx, dy2);
        const bool plo = below_edge<T>(A, mid_e, d2, L.xa, L.ya, q.x, q.y);
        T m = plo ? (T)1 : (T)0;
        T m2 = (T)1 - m;
        if (BOUNDED) {
            const bool ge_lo = (lo_e < 0) ? true : !below_edge<T>(A, lo_e, d2, L.xa, L.ya, q.x, q.y);
            const bool lt_hi = (hi_e > n_bin) ? true : below_edge<T>(A, hi_e, d2, L.xa, L.ya, q.x, q.y);
            m = (plo && ge_lo) ? (T)1 : (T)0;
            m2 = (!plo && lt_hi) ? (T)1 : (T)0;
        }
        if (KIND == SHEAR && ZCHECK) {
            if (L.za > zr[j]) { m = 0; m2 = 0; }   // source in front of the lens, src/tpcf.f90:322-324
        }
        T d, inv;
        Geo<T>::dinv(d2, d, inv);
        // keep every per-pair value finite so that 0/1 masks cannot leak a NaN
        // from an excluded pair; in-bin zero separations are tracked apart
        d = (d2 > (T)0) ? d : (T)0;
        inv = (d2 > (T)0) ? inv : (T)0;
        if (track0) sZ += (d2 == (T)0) ? m : (T)0;
        const T tv = d * q.w;
        sT0 = fma(tv, m, sT0);
        sT1 = fma(tv, m2, sT1);
        if (KIND == SHEAR) {
            const T u = fma(dx, dx, -dy2);
            const T v = dx * dy;
            const T g1 = fma(q.s1w, u, q.s2w2 * v) * inv;
            const T g2 = fma(q.s2w, u, -(q.s1w2 * v)) * inv;
            sG10 = fma(g1, m, sG10);
            sG11 = fma(g1, m2, sG11);
            sG20 = fma(g2, m, sG20);
            sG21 = fma(g2, m2, sG21);
        } else {
            sG10 = fma(q.s1w, m, sG10);
            sG11 = fma(q.s1w, m2, sG11);
            sG20 = fma(q.s2w, m, sG20);
            sG21 = fma(q.s2w, m2, sG21);
        }
        sW0 = fma(q.w, m, sW0);
        sW1 = fma(q.w, m2, sW1);
        sK0 = fma(q.wk, m, sK0);
        sK1 = fma(q.wk, m2, sK1);
        sN0 += m;
        sN1 += m2;
    }
    w.aT[0] += (double)sT0; w.aT[1] += (double)sT1;
    w.aG1[0] += (double)sG10; w.aG1[1] += (double)sG11;
    w.aG2[0] += (double)sG20; w.aG2[1] += (double)sG21;
    w.aW[0] += (double)sW0; w.aW[1] += (double)sW1;
    w.aK[0] += (double)sK0; w.aK[1] += (double)sK1;
    w.aN[0] += (double)sN0; w.aN[1] += (double)sN1;
    w.aZ += (double)sZ;
    w.used0 = w.used1 = true;
    w.edge = true;
}

template <typename T, int KIND, bool MARKS>
__global__ __launch_bounds__(kWave *kWavesPerBlock) void k_pairs(
    const PairArgs<T> A, const Chunk<T> *__restrict__ chunks, const T *__restrict__ lx, const T *__restrict__ ly,
    const T *__restrict__ lz, const T *__restrict__ la, const T *__restrict__ lw, const Seg<T> *__restrict__ segs,
    const uint32_t *__restrict__ seg_first, const int4 *__restrict__ pbox, const SrcRec<T> *__restrict__ srec,
    const T *__restrict__ sz, const T *__restrict__ e2lane, double *__restrict__ pp,
    unsigned long long *__restrict__ ppn, double *__restrict__ marks, unsigned long long *__restrict__ probe)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int64_t b = blockIdx.x;
    const int slot = (int)(b % kSlots);
    const int64_t ci = ((b % kSlots) * A.per_xcd + b / kSlots) * kWavesPerBlock + wv;
    if (ci >= A.n_chunks) return;
    const Chunk<T> ch = chunks[ci];
    LaneData<T> L;
    L.lane = lane;
    L.valid = lane < (int)ch.count;
    L.li = ch.start + (L.valid ? lane : 0);
    L.xa = lx[L.li];
    L.ya = ly[L.li];
    L.za = (KIND == SHEAR) ? lz[L.li] : (T)0;
    L.wa = L.valid ? lw[L.li] : (T)0;
    L.aa = L.valid ? la[L.li] : (T)0;
    L.vmask = L.valid ? (T)1 : (T)0;
    const int n_bin = A.n_bin;
    const T e2_lane = e2lane[lane];   // squared edge `lane` (+inf past n_bin)

    Win w;
    win_clear(w);
    w.pb = -1;
    w.wk = -3;
    Table t = {0, 0, 0, 0, 0, 0};

    const double xm = A.x_max, mg = A.margin;
    const Grid g = A.g;
    for (int pb = 0; pb < A.n_pb; ++pb) {
        const int4 pbx = pbox[pb];
        if (pbx.x > pbx.y) continue;   // no sources in this patch
        const int cy0 = max(cell_of((double)ch.by0 - xm - mg, g.y0, g.inv, g.ny), pbx.z);
        const int cy1 = min(cell_of((double)ch.by1 + xm + mg, g.y0, g.inv, g.ny), pbx.w);
        if (cy0 > cy1) continue;
        for (int cy = cy0; cy <= cy1; ++cy) {
            const double ry0 = g.y0 + (double)cy * g.cell - mg, ry1 = g.y0 + (double)(cy + 1) * g.cell + mg;
            const double dyl = fmax(0.0, fmax(ry0 - (double)ch.by1, (double)ch.by0 - ry1));
            if (dyl > xm) continue;
            const double hw = sqrt(fmax(xm * xm - dyl * dyl, 0.0)) + mg;
            const int cx0 = max(cell_of((double)ch.bx0 - hw, g.x0, g.inv, g.nx), pbx.x);
            const int cx1 = min(cell_of((double)ch.bx1 + hw, g.x0, g.inv, g.nx), pbx.y);
            if (cx0 > cx1) continue;
            const int64_t key0 = (int64_t)pb * g.ncells + (int64_t)cy * g.nx + cx0;
            const uint32_t s_begin = __builtin_amdgcn_readfirstlane(seg_first[key0]);
            const uint32_t s_end = __builtin_amdgcn_readfirstlane(seg_first[key0 + (cx1 - cx0) + 1]);
            for (uint32_t si = s_begin; si < s_end; ++si) {
                const Seg<T> sg = segs[si];
                const uint32_t cnt = __builtin_amdgcn_readfirstlane(sg.count);
                const SrcRec<T> *__restrict__ r = srec + __builtin_amdgcn_readfirstlane(sg.start);
                // squared distance range between the chunk box and the segment box
                const T dxl = fmax((T)0, fmax(sg.x0 - ch.bx1, ch.bx0 - sg.x1));
                const T dyl2 = fmax((T)0, fmax(sg.y0 - ch.by1, ch.by0 - sg.y1));
                const T dxh = fmax(sg.x1 - ch.bx0, ch.bx1 - sg.x0);
                const T dyh = fmax(sg.y1 - ch.by0, ch.by1 - sg.y0);
                const T lo2 = (dxl * dxl + dyl2 * dyl2) * A.cls_lo;
                const T hi2 = (dxh * dxh + dyh * dyh) * A.cls_hi;
                const int kmin = __popcll(__ballot(e2_lane <= lo2)) - 1;
                const int kmax = __popcll(__ballot(e2_lane <= hi2)) - 1;
                if (kmax < 0 || kmin >= n_bin) continue;   // all pairs outside [x_min, x_max]
                const bool zc = (KIND == SHEAR) && (ch.zmax > sg.zmin);
                int cls;
                if (kmin == kmax && !zc && lo2 > (T)0) {
                    cls = 0;
                    // ---- fast: the whole segment is in bin kmin ----
                    if (!(w.open && pb == w.pb && (kmin == w.wk || kmin == w.wk + 1)))
                        win_open<T, KIND, MARKS>(w, t, L, n_bin, marks, pb, kmin);
                    T sT = 0, sG1 = 0, sG2 = 0;
                    if (KIND == SHEAR) {
#pragma unroll 4
                        for (uint32_t j = 0; j < cnt; ++j) {
                            const SrcRec<T> q = r[j];
                            const T dx = L.xa - q.x, dy = L.ya - q.y;
                            const T dy2 = dy * dy;
                            const T d2 = fma(dx, dx, dy2);
                            const T u = fma(dx, dx, -dy2);   // d^2 cos(2 phi)
                            const T v = dx * dy;             // d^2 sin(2 phi) / 2
                            T d, inv;
                            Geo<T>::dinv(d2, d, inv);
                            sT = fma(d, q.w, sT);
                            sG1 = fma(fma(q.s1w, u, q.s2w2 * v), inv, sG1);
                            sG2 = fma(fma(q.s2w, u, -(q.s1w2 * v)), inv, sG2);
                        }
                    } else {
#pragma unroll 4
                        for (uint32_t j = 0; j < cnt; ++j) {
                            const SrcRec<T> q = r[j];
                            const T dx = L.xa - q.x, dy = L.ya - q.y;
                            const T d2 = fma(dx, dx, dy * dy);
                            sT = fma(Geo<T>::dist(d2), q.w, sT);
                        }
                    }
                    const int s = kmin - w.wk;
                    if (s == 0) {
                        w.aT[0] += (double)sT; w.aG1[0] += (double)sG1; w.aG2[0] += (double)sG2;
                        w.used0 = true;
                        if (MARKS) {
                            w.aW[0] += sg.W; w.aK[0] += sg.K;
                            if (KIND == KAPPA) { w.aG1[0] += sg.S1; w.aG2[0] += sg.S2; }
                        }
                    } else {
                        w.aT[1] += (double)sT; w.aG1[1] += (double)sG1; w.aG2[1] += (double)sG2;
                        w.used1 = true;
                        if (MARKS) {
                            w.aW[1] += sg.W; w.aK[1] += sg.K;
                            if (KIND == KAPPA) { w.aG1[1] += sg.S1; w.aG2[1] += sg.S2; }
                        }
                    }
                    // pair count (and, without marks, W, K and kappa's G) from the segment sums
                    if (lane == kmin) {
                        t.N += (double)cnt * (double)ch.count;
                        if (!MARKS) {
                            t.W += ch.wsum * sg.W;
                            t.K += ch.wsum * sg.K;
                            if (KIND == KAPPA) { t.G1 += ch.asum * sg.S1; t.G2 += ch.asum * sg.S2; }
                        }
                    }
                } else if (kmax == kmin + 1 && !zc) {
                    cls = 1;
                    // ---- edge: one bin edge inside the distance range ----
                    win_open<T, KIND, MARKS>(w, t, L, n_bin, marks, pb, kmin);
                    win_pass<T, KIND, false, false>(w, L, A, r, nullptr, cnt, n_bin);
                } else {
                    cls = 2;
                    // ---- multi: two-bin windows over [kmin, kmax], bounded tests ----
                    const T *zr = (KIND == SHEAR) ? sz + __builtin_amdgcn_readfirstlane(sg.start) : nullptr;
                    for (int w0 = kmin; w0 <= kmax; w0 += 2) {
                        win_open<T, KIND, MARKS>(w, t, L, n_bin, marks, pb, w0);
                        if (zc) win_pass<T, KIND, true, true>(w, L, A, r, zr, cnt, n_bin);
                        else win_pass<T, KIND, true, false>(w, L, A, r, zr, cnt, n_bin);
                        win_flush<T, KIND, MARKS>(w, t, L, n_bin, marks);
                    }
                }
                if (A.counters && lane == 0) {
                    atomicAdd(&probe[cls], 1ull);
                    atomicAdd(&probe[3 + cls], (unsigned long long)cnt * ch.count);
                }
            }
        }
        win_flush<T, KIND, MARKS>(w, t, L, n_bin, marks);
        // hand the patch's per-bin sums to the patch-pair accumulators
        if (lane < n_bin && t.N > 0) {
            const size_t o = (((size_t)slot * A.n_pa + ch.pa) * A.n_pb + pb) * n_bin + lane;
            unsafeAtomicAdd(&pp[o * kNQ + 0], t.T);
            unsafeAtomicAdd(&pp[o * kNQ + 1], t.G1);
            unsafeAtomicAdd(&pp[o * kNQ + 2], t.G2);
            unsafeAtomicAdd(&pp[o * kNQ + 3], t.W);
            unsafeAtomicAdd(&pp[o * kNQ + 4], t.K);
            atomicAdd(&ppn[o], (unsigned long long)t.N);
        }
        t = Table{0, 0, 0, 0, 0, 0};
    }
}

template <typename T>
int run_pairs(tpcf_ctx *ctx, const CallArgs<T> &a, const Prepared<T> &p, double *pp, unsigned long long *ppn,
              double *marks)
{
    PairArgs<T> A;
    std::memset(&A, 0, sizeof A);
    A.g = p.grid;
    A.n_chunks = p.n_chunks;
    const int64_t nblk = (p.n_chunks + kWavesPerBlock - 1) / kWavesPerBlock;
    A.per_xcd = (nblk + kSlots - 1) / kSlots;
    A.n_pa = a.n_pa;
    A.n_pb = a.n_pb;
    A.n_bin = a.bins->n_bin;
    A.counters = ctx->opt_counters ? 1 : 0;
    const int n = A.n_bin;
    A.x_max = a.bins->edges[n];
    A.xmin_zero = (a.bins->edges[0] == 0.0) ? 1 : 0;
    const double scale = std::fabs(p.grid.x0) + std::fabs(p.grid.y0) + p.grid.cell * (p.grid.nx + p.grid.ny) + A.x_max;
    A.margin = 1e-9 * scale + 1e-12;
    if (sizeof(T) == 4) {
        A.cls_lo = (T)(1.0 - 2e-6);
        A.cls_hi = (T)(1.0 + 2e-6);
    } else {
        A.cls_lo = (T)(1.0 - 1e-12);
        A.cls_hi = (T)(1.0 + 1e-12);
    }
    A.band = 1e-5f;
    for (int i = 0; i <= TPCF_MAX_BIN; ++i) {
        if (i <= n) {
            const double e = a.bins->edges[i];
            A.e2d[i] = e * e;
            A.e2f[i] = (float)(e * e);
        } else {
            A.e2d[i] = INFINITY;
            A.e2f[i] = INFINITY;
        }
    }
    A.e2d[n] = std::nextafter(A.e2d[n], INFINITY);   // d == x_max belongs to the last bin

    T *e2lane;
    TPCF_TRY(ensure(ctx, tpcf_ctx::B_E2, sizeof(T) * kWave, (void **)&e2lane));
    {
        T h[kWave];
        for (int i = 0; i < kWave; ++i) h[i] = (i <= n) ? (sizeof(T) == 4 ? (T)A.e2f[i] : (T)A.e2d[i]) : (T)INFINITY;
        TPCF_HIP(ctx, hipMemcpyAsync(e2lane, h, sizeof h, hipMemcpyHostToDevice, ctx->stream));
        TPCF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    unsigned long long *probe = nullptr;
    if (A.counters) {
        TPCF_TRY(ensure(ctx, tpcf_ctx::B_PROBE, sizeof(unsigned long long) * 8, (void **)&probe));
        fill_u64(ctx->stream, probe, 8);
    }
    if (p.n_chunks == 0) return TPCF_OK;
    const dim3 grid((unsigned)(A.per_xcd * kSlots)), block(kWave * kWavesPerBlock);
    TPCF_HIP(ctx, hipEventRecord(ctx->ev[2], ctx->stream));
#define TPCF_LAUNCH(K, M)                                                                                         \
    hipLaunchKernelGGL((k_pairs<T, K, M>), grid, block, 0, ctx->stream, A, p.chunks, p.lx, p.ly, p.lz, p.la, p.lw, \
                       p.segs, p.seg_first, p.pbox, p.srec, p.sz, e2lane, pp, ppn, marks, probe)
    if (a.kind == SHEAR) {
        if (a.marks) TPCF_LAUNCH(SHEAR, true); else TPCF_LAUNCH(SHEAR, false);
    } else {
        if (a.marks) TPCF_LAUNCH(KAPPA, true); else TPCF_LAUNCH(KAPPA, false);
    }
#undef TPCF_LAUNCH
    TPCF_HIP(ctx, hipGetLastError());
    TPCF_HIP(ctx, hipEventRecord(ctx->ev[3], ctx->stream));
    return TPCF_OK;
}

template int run_pairs<float>(tpcf_ctx *, const CallArgs<float> &, const Prepared<float> &, double *,
                              unsigned long long *, double *);
template int run_pairs<double>(tpcf_ctx *, const CallArgs<double> &, const Prepared<double> &, double *,
                               unsigned long long *, double *);

}  // namespace tpcf
