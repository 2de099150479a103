// pairs.hip -- the pair kernel: the inner loops of tangential_shear_tpcf
// (src/tpcf.f90:266-359) and scalar_kappa_tpcf (src/tpcf.f90:456-526).
//
// One wavefront owns one work item: a lens chunk (up to 64 lenses, one per
// lane) and one band of the cell rows it can reach.  It walks the source
// segments within x_max of the chunk's bounding box, source patch by source
// patch, row by row.  Source records are
// wave-uniform, so they are read with scalar loads and every lane pairs its
// lens with the same source.  Each (chunk, segment) pair is classified from
// the two bounding boxes:
//   fast  -- every pair lies in one bin: no per-pair binning; W, K and the
//            pair count come from the segment sums (and G for kappa);
//   edge  -- pairs straddle one bin edge: per-pair FP32 test against the
//            squared edge, FP64 re-test for pairs within the band; each
//            side's sums are taken directly, in exec-masked branches;
//   multi -- several edges (or a z test): sweeps of up to 4 bins, each pair
//            added to its bin's lane-private row in LDS.
// Each segment's pairs are summed per lane in the input precision (at most
// kSegMax terms), then added to per-lane FP64 sums for a window of two
// adjacent bins (fast, edge) or reduced straight from the sweep's rows
// (multi).  When the window moves these are reduced across the wave
// (through LDS rows) into FP64 per-bin registers (lane b holds bin b), which
// go to the per-patch-pair FP64 accumulators with global atomics when the
// source patch changes.  On the sphere (SPH) the same passes run on FP64
// unit vectors: squared chords against squared chord edges, the position
// angle from the east and north components at the lens.  Pair counts are
// integers throughout.
#include <algorithm>
#include <cmath>
#include <type_traits>

#include "internal.hpp"

namespace tpcf {

void fill_u64(hipStream_t s, unsigned long long *p, int64_t n);
void fill_f64(hipStream_t s, double *p, int64_t n);


// 1/sqrt(d2) for d2 > 0 from the hardware reciprocal square root and one
// Newton-Raphson step r <- r (3/2 - d2 r^2 / 2), which takes the seed's
// relative error e to about 3 e^2 / 2: seed within 5.3e-8, result within
// 4.2e-15 (measured over 2^20 values spanning 2^-40 .. 2^5, against long
// double).  Used for separations and projection weights only, never for a
// bin decision (those compare squared distances).
__device__ __forceinline__ double rsq_pos(double d2)
{
    const double r = __builtin_amdgcn_rsq(d2);
    return r * fma(-(0.5 * d2) * r, r, 1.5);
}

__device__ __forceinline__ double sqrt_pos(double d2) { return d2 * rsq_pos(d2); }

// sqrt(d2) for d2 >= 0 from the hardware reciprocal square root r of d2 +
// 2^-1022 and one Newton-Raphson step on the root itself, s0 = d2 r,
// s = s0 + (r/2) (d2 - s0^2): a seed error e leaves 3 e^2 / 2, as rsq_pos,
// in one instruction fewer than d2 rsq_pos(d2).  d2 + 2^-1022 is d2 itself
// for every d2 >= 2^-969, and d2 == 0 gives 0 without a compare and a
// select.
__device__ __forceinline__ double sqrt_nr(double d2)
{
    const double r = __builtin_amdgcn_rsq(d2 + 0x1p-1022);
    const double s0 = d2 * r;
    return fma(0.5 * r, fma(-s0, s0, d2), s0);
}

// 1/s for s > 0 from the hardware reciprocal and one Newton-Raphson step
// r <- r + r (1 - s r): seed within 4.7e-8, result within 2.3e-15 (measured
// as for rsq_pos).
__device__ __forceinline__ double rcp_pos(double s)
{
    const double r = __builtin_amdgcn_rcp(s);
    return fma(r, fma(-s, r, 1.0), r);
}

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
    // separation and 1/d^2 from d^2 > 0 with one refined reciprocal square
    // root (the IEEE-rounded sqrt and division sequences cost twice as much)
    static __device__ inline void dinv(double d2, double &d, double &inv)
    {
        const double r = rsq_pos(d2);
        d = d2 * r;
        inv = r * r;
    }
    static __device__ inline double dist(double d2) { return sqrt_nr(d2); }
};

// d^2 < edge test for the squared edge with index e.  FP32: decided in FP32
// unless d^2 lies within the band [lo, hi] = edge^2 (1 -/+ band) around the
// edge, then re-decided in FP64 from the FP32 inputs exactly as the FP64
// oracle does.  Outside the band d2 < lo and d2 < edge^2 agree, so two
// compares (d2 < lo, d2 <= hi) give both the FP32 decision and the band.
// FP64: direct.
template <typename T>
struct EdgeTest {
    float lo, hi;
    double ed;
    __device__ __forceinline__ EdgeTest(const PairArgs<T> &A, int e)
        : lo(A.e2f[e] - A.band * A.e2f[e]), hi(A.e2f[e] + A.band * A.e2f[e]), ed(A.e2d[e]) {}
    __device__ __forceinline__ bool below(T d2, T xa, T ya, T xs, T ys) const
    {
        if constexpr (sizeof(T) == 4) {
            bool p = d2 < lo;
            if (!p && d2 <= hi) {
                asm volatile("" : "+v"(xa), "+v"(ya));
                const double dx = (double)xa - (double)xs, dy = (double)ya - (double)ys;
                p = fma(dx, dx, dy * dy) < ed;
            }
            return p;
        } else {
            return d2 < ed;
        }
    }
};

// Makes the record's values inputs of an (empty) instruction: the wait for
// its scalar loads is placed here, before the loads issued next.
template <typename T>
__device__ __forceinline__ void arrived(const SrcRec<T> &q)
{
    if constexpr (sizeof(T) == 4)
        asm volatile("" ::"s"(q.x), "s"(q.y), "s"(q.w), "s"(q.wk), "s"(q.s1w), "s"(q.s2w), "s"(q.s1w2), "s"(q.s2w2));
    else
        asm volatile("" ::"s"(q.x), "s"(q.y), "s"(q.w), "s"(q.wk), "s"(q.s1w), "s"(q.s2w), "s"(q.s1w2), "s"(q.s2w2));
}

// Visits r[0..count) two records at a time while the next two are already
// requested, alternating between two register sets (no copies).  Scalar loads
// return out of order, so a wait covers every load in flight: each set is
// waited for (`arrived`) before the loads of the other set are issued.  The
// record array is padded with kRecPad records, so reading up to three records
// past a segment's end stays inside the allocation.
// The source records are read through a pointer into the constant address
// space: loads from it are scalar whenever the address is wave-uniform (a
// global pointer that moves through the loop loses the provenance that lets
// the compiler prove the records unmodified, and is then read with vector
// loads).  The records are written only by the preparation kernels.
template <typename T>
using ConstRec = const __attribute__((address_space(4))) SrcRec<T>;

template <typename T>
__device__ __forceinline__ ConstRec<T> *const_rec(const SrcRec<T> *r)
{
    return (ConstRec<T> *)r;
}

template <typename T>
__device__ __forceinline__ SrcRec<T> rec_at(ConstRec<T> *r)
{
    SrcRec<T> q;
    q.x = r->x; q.y = r->y; q.w = r->w; q.wk = r->wk;
    q.s1w = r->s1w; q.s2w = r->s2w; q.s1w2 = r->s1w2; q.s2w2 = r->s2w2;
    return q;
}

template <typename T, typename F>
__device__ __forceinline__ void stream2(const SrcRec<T> *__restrict__ rg, uint32_t count, F &&f)
{
    if (count == 0) return;
    // a moving record pointer and a down-counter: both change every half
    // iteration, so they stay in registers (a fixed base pointer live across
    // the loop was displaced and reloaded for every load)
    ConstRec<T> *r = const_rec(rg);
    SrcRec<T> a0 = rec_at(r), a1 = rec_at(r + 1);
    uint32_t left = count;
    for (;;) {
        arrived(a0);
        arrived(a1);
        const SrcRec<T> b0 = rec_at(r + 2), b1 = rec_at(r + 3);
        asm volatile("" ::: "memory");   // the loads stay ahead of this half's work
        f(a0);
        if (left > 1) f(a1);
        if (left <= 2) break;
        left -= 2;
        r += 2;
        arrived(b0);
        arrived(b1);
        a0 = rec_at(r + 2);
        a1 = rec_at(r + 3);
        asm volatile("" ::: "memory");
        f(b0);
        if (left > 1) f(b1);
        if (left <= 2) break;
        left -= 2;
        r += 2;
    }
}

// Per-wave window state: two adjacent bins (wk, wk+1) of one source patch.
// The per-lane FP64 window sums live in LDS, lane-private: acc[q][lane] for
// q = Theta, G1, G2, W, K of bin wk (q) and bin wk+1 (WQ + q), plus the
// count of in-bin zero-separation pairs (shear with x_min == 0).  Pair counts
// are wave totals (ballot + popcount) and go straight to the per-bin sums,
// q = Theta, G1, G2, W, K kept in FP64 registers and the pair count N in a
// u64 register, by the lane of the bin (K only when some m != 0: otherwise it
// equals W and is not kept).  Only
// the window sums (or, for a multi segment, the sweep's rows in their place)
// and the grid use LDS: 4.6 KB per wave when all m == 0, 5.7 KB otherwise,
// which leaves room for 8 and 7 workgroups of 4 waves per CU.
// window quantities per bin: Theta, G1, G2, W (+ K unless all m == 0)
template <bool MZ>
constexpr int win_q() { return MZ ? 4 : 5; }
enum { QT = 0, QG1 = 1, QG2 = 2, QW = 3, QK = 4 };

struct Win {
    int pb, wk;
    bool open, used0, used1, edge;
    double *acc;   // this wave's [2 WQ + 1][kWave] block in LDS
    int64_t mstride;   // n_lens: the stride of the marks planes
    double t[kNQ];   // per-bin sums Theta, G1, G2, W, K of bin `lane`, in registers (lane b owns bin b)
    uint64_t tn;     // pair count of bin `lane` (integer: exact at any count)
    uint64_t valid;   // lanes holding a lens
    uint32_t n_valid;
};

// Adds a wave-total pair count to the table entry of bin `bin` (if real).
__device__ __forceinline__ void tab_count(Win &w, int lane, int bin, int n_bin, uint64_t n)
{
    if (bin >= 0 && bin < n_bin && lane == bin) w.tn += n;
}

__device__ __forceinline__ double &wacc(const Win &w, int q, int lane) { return w.acc[q * kWave + lane]; }

template <typename T>
struct LaneData {
    T xa, ya, za, wa, aa, vmask;
    bool valid;
    uint32_t li;
    int lane;
    // spherical geometry: third rotated component and the east / north unit
    // vectors at the lens in the rotated basis (p, q, s)
    T pa, fe0, fe1, fe2, fn0, fn1, fn2;
};

// Pair geometry.  Flat sky: (a, b) = (x_A - x_B, y_A - y_B), c = b^2,
// d2 = a^2 + b^2; the projection terms are u = a^2 - b^2 = d^2 cos 2phi,
// v = a b = d^2 sin 2phi / 2 and inv = 1/d^2 (src/tpcf.f90:317-319,
// :343-344).  Sphere: (a, b, c) = u_B - u_A in the rotated basis (the
// source's p sits in record slot s1w2), d2 = |u_B - u_A|^2 (the squared
// chord); E, N = components along the east and north unit vectors at the
// lens, u = E^2 - N^2, v = E N, inv = 1/(E^2 + N^2) = 1/sin^2 d, which are the
// reference's cos 2phi and sin 2phi from angles_sphere (src/utils.f90:618-639)
// without trigonometry; the separation angle is 2 asin(chord / 2).
// sph_rev(SPH) is the sphere with the roles of the catalogues swapped (the
// adjoint's source-side pass: the lanes hold sources, the stream holds
// lenses) while the angle stays the one at the lens, i.e. at the streamed
// object A.  With P the pole and D = u_A - u_B, the direction at A toward B
// scaled by cos(lat_A) > 0 is E' = D . (u_B x P), N' = -P . D + (P . u_A) |D|^2 / 2
// (from e_A = P x u_A / cos lat_A, n_A = u_A x e_A); the lane's frame slots
// hold u_B x P and P, and pu = P . u_A.  The common factor drops out of u, v
// and inv.
// SPH: 0 the flat sky, 1 the sphere, 2 the sphere with the catalogues' roles
// swapped; plus 4 where every chord of the segment visit is below kSphShort
// (the pair kernel picks that variant per visit).
constexpr bool sph_rev(int sph) { return (sph & 3) == 2; }
constexpr bool sph_short(int sph) { return (sph & 4) != 0; }

template <typename T, int SPH>
struct PairGeo {
    T a, b, c, d2;
    T pu;   // sph_rev(SPH): P . u_A of the streamed object
};

template <typename T, int SPH>
__device__ __forceinline__ PairGeo<T, SPH> pair_geo(const LaneData<T> &L, const SrcRec<T> &q)
{
    PairGeo<T, SPH> g;
    if constexpr (!SPH) {
        g.a = L.xa - q.x;
        g.b = L.ya - q.y;
        g.c = g.b * g.b;
        g.d2 = fma(g.a, g.a, g.c);
    } else {
        g.a = q.s1w2 - L.pa;
        g.b = q.x - L.xa;
        g.c = q.y - L.ya;
        g.d2 = fma(g.a, g.a, fma(g.b, g.b, g.c * g.c));
        if constexpr (sph_rev(SPH)) g.pu = fma(L.fn0, q.s1w2, fma(L.fn1, q.x, L.fn2 * q.y));
    }
    return g;
}

// asin(x)/x = sum_k C(2k,k) x^2k / (4^k (2k+1)) to x^12, for x^2 = x2 <
// 0.025: the first omitted term, 143/10240 x^14, is below 1e-16 of the sum
// for x2 < 9.5e-3 and below 8.6e-14 at 0.025.
__device__ __forceinline__ double asin_ratio(double x2)
{
    double t = 231.0 / 13312.0;
    t = fma(t, x2, 63.0 / 2816.0);
    t = fma(t, x2, 35.0 / 1152.0);
    t = fma(t, x2, 5.0 / 112.0);
    t = fma(t, x2, 3.0 / 40.0);
    t = fma(t, x2, 1.0 / 6.0);
    return fma(t, x2, 1.0);
}

// a b + c in one VOP3 FMA, with the constant c read from a scalar register
// pair.  Left to itself the compiler keeps an FP64 constant in vector
// registers and forms the two-address FMA whose accumulator is the constant,
// which costs a register copy for every term of a series.
__device__ __forceinline__ double fma_sc(double a, double b, double c)
{
    double r;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "s"(c));
    return r;
}

// asin_ratio(d2 / 4) as a series in d2 itself: the coefficients divided by
// 4^k, which changes no bit (every Horner step is the one of asin_ratio
// scaled by a power of two), and the quartering of d2 is saved.  SHORT: the
// terms to x^8 only, for d2 < kSphShort, where the first omitted term,
// 63/2816 (d2/4)^5, is at most 1.0e-16 of the sum.
constexpr double kSphShort = 5.4e-3;

template <bool SHORT>
__device__ __forceinline__ double asin_ratio_c2(double d2)
{
    double t;
    if constexpr (SHORT) {
        t = fma_sc(d2, 35.0 / 1152.0 / 256.0, 5.0 / 112.0 / 64.0);
    } else {
        t = fma_sc(d2, 231.0 / 13312.0 / 4096.0, 63.0 / 2816.0 / 1024.0);
        t = fma_sc(t, d2, 35.0 / 1152.0 / 256.0);
        t = fma_sc(t, d2, 5.0 / 112.0 / 64.0);
    }
    t = fma_sc(t, d2, 3.0 / 40.0 / 16.0);
    t = fma_sc(t, d2, 1.0 / 6.0 / 4.0);
    return fma(t, d2, 1.0);
}

// 2 asin(x) for x^2 = x2 in [0.025, 1] (18 to 180 degrees): the half-angle
// step asin(x) = 2 asin(x'), x'^2 = x^2 / (2 (1 + sqrt(1 - x^2))), taken until
// x'^2 < 0.025 (at most four times: sin^2(pi/32) < 0.025), then the series.
// It needs a few registers, where an inlined library asin raised the pair
// loops' register peak and made them spill.
__device__ __noinline__ double sph_angle_wide(double x2)
{
    double scale = 2.0;
    for (int k = 0; k < 4 && x2 >= 0.025; ++k) {
        const double r = 1.0 - x2;
        const double s = r > 0.0 ? sqrt_pos(r) : 0.0;
        x2 = x2 * rcp_pos(2.0 + 2.0 * s);
        scale += scale;
    }
    return scale * sqrt_pos(x2) * asin_ratio(x2);
}

// Separation angle from the squared chord c^2 = d2: 2 asin(x) with x = c/2,
// = c asin(x)/x; the series below x^2 = d2/4 = 0.025 (18 degrees).  A zero
// chord gives 0 (sqrt_nr).  SHORT: d2 < kSphShort is known for every pair.
template <bool SHORT>
__device__ __forceinline__ double sph_angle(double d2)
{
    if constexpr (SHORT) return sqrt_nr(d2) * asin_ratio_c2<true>(d2);
    if (d2 < 0.1) return sqrt_nr(d2) * asin_ratio_c2<false>(d2);
    return sph_angle_wide(fmin(0.25 * d2, 1.0));
}

// Separation, 1/d^2 (or 1/sin^2 d) and the projection terms of one pair.  At
// zero separation d, u and v are 0, and so is inv on the flat sky with GUARD
// (on the sphere inv stays finite there, and inv u, inv v are 0).
template <typename T, int SPH, bool GUARD>
__device__ __forceinline__ void pair_proj(const LaneData<T> &L, const PairGeo<T, SPH> &g, T &d, T &inv, T &u, T &v)
{
    if constexpr (!SPH) {
        Geo<T>::dinv(g.d2, d, inv);
        if (GUARD) {
            d = (g.d2 > (T)0) ? d : (T)0;
            inv = (g.d2 > (T)0) ? inv : (T)0;
        }
        u = fma(g.a, g.a, -g.c);
        v = g.a * g.b;
    } else {
        const T E = fma(L.fe0, g.a, fma(L.fe1, g.b, L.fe2 * g.c));
        T N = fma(L.fn0, g.a, fma(L.fn1, g.b, L.fn2 * g.c));
        if constexpr (sph_rev(SPH)) N = fma((T)0.5 * g.d2, g.pu, -N);
        const T s2 = fma(E, E, N * N);
        // s2 == 0 (zero separation, or E and N below 1e-154) leaves u = v =
        // 0, so the floor at the smallest normal number keeps the pair's
        // terms at 0 without a compare and a select
        inv = rcp_pos(fmax(s2, 0x1p-1022));
        u = fma(E, E, -(N * N));
        v = E * N;
        d = sph_angle<sph_short(SPH)>(g.d2);
    }
}

// 2 e1 w of a record (slot s1w2 holds p on the sphere)
template <typename T, int SPH>
__device__ __forceinline__ T s1w2_of(const SrcRec<T> &q)
{
    if constexpr (SPH) return q.s1w + q.s1w;
    else return q.s1w2;
}

template <bool MZ>
__device__ __forceinline__ void win_clear(Win &w, int lane)
{
#pragma unroll
    for (int q = 0; q < 2 * win_q<MZ>() + 1; ++q) wacc(w, q, lane) = 0.0;
    w.open = false;
    w.used0 = w.used1 = false;
    w.edge = false;
}

// Sums over the wave of rows r < NR <= 8 of 64 per-lane FP64 values
// (rows[r * kWave + lane]) through LDS rather than cross-lane steps (a DPP
// reduction of one FP64 value takes about 20 VALU instructions): lane l adds
// the values 8 (l % 8) .. 8 (l % 8) + 7 of row l / 8 and writes the partial to
// slot l % 8 of that row, then lanes r < NR add their row's 8 partials and
// leave the total in slot 8.  The rows are overwritten.  A wave's LDS
// operations run in order, so a step reads what the step before wrote
// without a barrier; all lanes of the wave must be active.
template <int NR>
__device__ __forceinline__ void lds_rowsum(double *rows, int lane)
{
    static_assert(NR >= 1 && NR <= 8, "eight lanes per row");
    const int r = lane >> 3, p = lane & 7;
    if (r < NR) {
        const double2 *q = reinterpret_cast<const double2 *>(rows + r * kWave + p * 8);
        const double2 a = q[0], b = q[1], c = q[2], d = q[3];
        rows[r * kWave + p] = ((a.x + a.y) + (b.x + b.y)) + ((c.x + c.y) + (d.x + d.y));
    }
    if (lane < NR) {
        const double2 *q = reinterpret_cast<const double2 *>(rows + lane * kWave);
        const double2 a = q[0], b = q[1], c = q[2], d = q[3];
        rows[lane * kWave + 8] = ((a.x + a.y) + (b.x + b.y)) + ((c.x + c.y) + (d.x + d.y));
    }
}

// Reduce the window's per-lane sums across the wave into the per-bin table
// (and, with marks, add each lane's sums to its lens' marks).  Each bin's
// rows are replaced by the lanes' weighted terms and summed by lds_rowsum.
template <typename T, int KIND, bool MARKS, bool MZ>
__device__ __forceinline__ void win_flush(Win &w, const LaneData<T> &L, int n_bin, double *__restrict__ marks)
{
    constexpr int WQ = win_q<MZ>();
    if (!w.open) return;
    const int lane = L.lane;
    const double dwa = (double)L.wa, daa = (double)L.aa;
    const double gs = (KIND == SHEAR) ? -1.0 : 1.0;   // G_t -= ... for shear, src/tpcf.f90:348-349
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const int bin = w.wk + s;
        const bool used = s ? w.used1 : w.used0;
        if (!used || bin < 0 || bin >= n_bin) continue;
        const int o = s * WQ;
        const double aT = wacc(w, o + QT, lane), aG1 = wacc(w, o + QG1, lane), aG2 = wacc(w, o + QG2, lane);
        const bool with_g = KIND == SHEAR || MARKS || w.edge, with_w = MARKS || w.edge;
        double aW = 0, aK = 0;
        if (with_w) {
            aW = wacc(w, o + QW, lane);
            aK = MZ ? aW : wacc(w, o + QK, lane);   // m == 0: K == W
        }
        // the position angle of a zero-separation pair is 0/0: the reference's
        // G_t and G_x of that bin, and the xi_E / xi_B marks of that lens,
        // become NaN (src/tpcf.f90:343-355)
        const bool zero_sep = KIND == SHEAR && bin == 0 && w.edge && (double)L.vmask * wacc(w, 2 * WQ, lane) > 0.0;
        wacc(w, o + QT, lane) = dwa * aT;
        wacc(w, o + QG1, lane) = with_g ? daa * aG1 : 0.0;
        wacc(w, o + QG2, lane) = with_g ? daa * aG2 : 0.0;
        wacc(w, o + QW, lane) = with_w ? dwa * aW : 0.0;
        if (!MZ) wacc(w, o + QK, lane) = with_w ? dwa * aK : 0.0;
        lds_rowsum<WQ>(w.acc + o * kWave, lane);
        if (lane == bin) {
            w.t[QT] += wacc(w, o + QT, 8);
            w.t[QG1] += gs * wacc(w, o + QG1, 8);
            w.t[QG2] += gs * wacc(w, o + QG2, 8);
            w.t[QW] += wacc(w, o + QW, 8);
            if (!MZ) w.t[QK] += wacc(w, o + QK, 8);   // MZ: K == W, not kept
        }
        if (KIND == SHEAR && bin == 0 && w.edge && __ballot(zero_sep) && lane == bin) {
            w.t[QG1] = __builtin_nan("");
            w.t[QG2] = __builtin_nan("");
        }
        // A lens' chunk may be split over several waves: add atomically.  The
        // atomics are what a marks pass costs, so a lane whose four terms are
        // all zero (no source with weight in this bin, or a masked lens) adds
        // nothing (the marks start at +0 and adding zeros leaves them as they
        // are), and with all m == 0 the K marks equal the W marks and are
        // copied from them after the kernel.
        if (MARKS && L.valid) {
            const double nan_or_0 = zero_sep ? __builtin_nan("") : 0.0;
            const double m0 = gs * daa * aG1 + nan_or_0, m1 = gs * daa * aG2 + nan_or_0;
            const double m2 = dwa * aK, m3 = dwa * aW;
            if (m0 != 0.0 || m1 != 0.0 || m2 != 0.0 || m3 != 0.0) {   // NaN != 0
                const int64_t n = w.mstride;
                double *mk = marks + mark_at(n, bin, 0, L.li);
                unsafeAtomicAdd(&mk[0], m0);
                unsafeAtomicAdd(&mk[n], m1);
                if (!MZ) unsafeAtomicAdd(&mk[2 * n], m2);
                unsafeAtomicAdd(&mk[3 * n], m3);
            }
        }
    }
    win_clear<MZ>(w, lane);
}

template <typename T, int KIND, bool MARKS, bool MZ>
__device__ __forceinline__ void win_open(Win &w, const LaneData<T> &L, int n_bin, double *marks, int pb, int k)
{
    if (w.open && (pb != w.pb || k != w.wk)) win_flush<T, KIND, MARKS, MZ>(w, L, n_bin, marks);
    w.pb = pb;
    w.wk = k;
    w.open = true;
}

// The edge decision of the edge passes as a lane mask (the lanes whose pair
// lies below the edge): FP32 d2 < lo, and for the lanes inside the band
// [lo, hi] the FP64 decision, in a branch the wave takes only when such a
// lane exists.  The masks stay in scalar registers: a lane's side is read back
// with inverse_ballot where the sums are taken.
template <typename T>
__device__ __forceinline__ uint64_t edge_below(const EdgeTest<T> &et, T d2, const LaneData<T> &L, const SrcRec<T> &q)
{
    if constexpr (sizeof(T) == 4) {
        uint64_t below = __builtin_amdgcn_ballot_w64(d2 < et.lo);
        const uint64_t amb = __builtin_amdgcn_ballot_w64(d2 <= et.hi) & ~below;
        if (amb) {
            // opaque copies: the FP64 lens coordinates are formed here,
            // not hoisted out of the loop into four long-lived registers
            float xa = L.xa, ya = L.ya;
            asm volatile("" : "+v"(xa), "+v"(ya));
            const double ddx = (double)xa - (double)q.x, ddy = (double)ya - (double)q.y;
            below |= __builtin_amdgcn_ballot_w64(fma(ddx, ddx, ddy * ddy) < et.ed) & amb;
        }
        return below;
    } else {
        return __builtin_amdgcn_ballot_w64(d2 < et.ed);
    }
}

// Edge pass: the segment straddles one edge (between window bins wk and
// wk+1).  Every sum is taken per side of the edge, directly: a side derived
// as total - other side would carry the other side's rounding, which is
// large next to a side that holds few of the segment's pairs.  The two sides
// are two exec-masked branches (empty asm statements keep them from being
// turned into selects), so each pair issues one side's instructions.  Pair
// counts are per lane, added in the upper side's branch, and summed over the
// valid lanes after the segment (a scalar count per source, the popcount of
// the lane mask, cost more: the kernel's scalar registers are its scarcest).  GUARD: the chunk and segment
// boxes overlap, so a pair may sit at zero separation (its per-pair values are
// forced finite); TRACK0: shear with x_min == 0, where such a pair is in bin 0
// (reference: NaN).  MZ: all m == 0, so K == W.
template <typename T, int KIND, bool MZ, bool GUARD, bool TRACK0, int SPH>
__device__ __forceinline__ void edge_pass(Win &w, const LaneData<T> &L, const PairArgs<T> &A,
                                          const SrcRec<T> *__restrict__ r, uint32_t count, int n_bin)
{
    constexpr int WQ = win_q<MZ>();
    const EdgeTest<T> et(A, w.wk + 1);
    T lT = 0, hT = 0, lW = 0, hW = 0, lK = 0, hK = 0, lG1 = 0, lG2 = 0, hG1 = 0, hG2 = 0, sZ = 0;
    uint32_t hN = 0;   // this lane's pairs at or above the edge
    stream2<T>(r, count, [&](const SrcRec<T> &q) __attribute__((always_inline)) {
        const PairGeo<T, SPH> g = pair_geo<T, SPH>(L, q);
        const T d2 = g.d2;
        const bool below = __builtin_amdgcn_inverse_ballot_w64(edge_below<T>(et, d2, L, q));
        T d, inv, u, v;
        pair_proj<T, SPH, GUARD>(L, g, d, inv, u, v);
        // the shear terms before the branches (both sides use them); 1/d^2
        // is applied by the branches' FMAs
        T t1 = 0, t2 = 0;
        if (KIND == SHEAR) {
            t1 = fma(q.s1w, u, q.s2w2 * v);
            t2 = fma(q.s2w, u, -(s1w2_of<T, SPH>(q) * v));
        }
        if (below) {
            lT = fma(d, q.w, lT);
            lW += q.w;
            if (!MZ) lK += q.wk;
            if (KIND == SHEAR) {
                lG1 = fma(t1, inv, lG1);
                lG2 = fma(t2, inv, lG2);
                if (TRACK0) sZ += (d2 == (T)0) ? (T)1 : (T)0;
            } else {
                lG1 += q.s1w;
                lG2 += q.s2w;
            }
            asm volatile("");
        } else {
            ++hN;
            hT = fma(d, q.w, hT);
            hW += q.w;
            if (!MZ) hK += q.wk;
            if (KIND == SHEAR) {
                hG1 = fma(t1, inv, hG1);
                hG2 = fma(t2, inv, hG2);
            } else {
                hG1 += q.s1w;
                hG2 += q.s2w;
            }
            asm volatile("");
        }
    });
    const int lane = L.lane;
    wacc(w, QT, lane) += (double)lT;
    wacc(w, WQ + QT, lane) += (double)hT;
    wacc(w, QG1, lane) += (double)lG1;
    wacc(w, QG2, lane) += (double)lG2;
    wacc(w, WQ + QG1, lane) += (double)hG1;
    wacc(w, WQ + QG2, lane) += (double)hG2;
    wacc(w, QW, lane) += (double)lW;
    wacc(w, WQ + QW, lane) += (double)hW;
    if (!MZ) {
        wacc(w, QK, lane) += (double)lK;
        wacc(w, WQ + QK, lane) += (double)hK;
    }
    if (KIND == SHEAR && TRACK0) wacc(w, 2 * WQ, lane) += (double)sZ;
    const uint32_t n_hi = wave_sum_u32(L.valid ? hN : 0u);   // the upper bin's pairs (a lens' lane must be valid)
    tab_count(w, lane, w.wk, n_bin, (uint64_t)count * w.n_valid - n_hi);
    tab_count(w, lane, w.wk + 1, n_bin, n_hi);
    w.used0 = w.used1 = true;
    w.edge = true;
}

// Edge pass at a boundary of the binned range: the segment straddles x_min
// (UPPER: of the window bins only wk+1 = 0 is in range) or x_max (!UPPER:
// only wk = n_bin - 1).  Only the in-range side is summed, directly; the
// values of out-of-range pairs, zero separations among them, are selected
// away.  Not used where a zero separation is in range (x_min == 0).  The
// kernel uses it at x_max, where the edge pass would give bin n_bin - 1 as
// the segment's G total minus the pairs beyond x_max, whose rounding can be
// large next to a bin that holds few of the segment's pairs.
template <typename T, int KIND, bool MZ, bool UPPER, int SPH>
__device__ __forceinline__ void edge_side_pass(Win &w, const LaneData<T> &L, const PairArgs<T> &A,
                                               const SrcRec<T> *__restrict__ r, uint32_t count, int n_bin)
{
    constexpr int WQ = win_q<MZ>();
    const EdgeTest<T> et(A, w.wk + 1);
    T sT = 0, sG1 = 0, sG2 = 0, sW = 0, sK = 0;
    uint32_t nIn = 0;   // this lane's in-range pairs
    stream2<T>(r, count, [&](const SrcRec<T> &q) __attribute__((always_inline)) {
        const PairGeo<T, SPH> g = pair_geo<T, SPH>(L, q);
        const uint64_t below = edge_below<T>(et, g.d2, L, q);
        const bool in = __builtin_amdgcn_inverse_ballot_w64(UPPER ? ~below : below);
        T d, inv, u, v;
        pair_proj<T, SPH, false>(L, g, d, inv, u, v);
        // an exec-masked branch (kept one by the empty asm): out-of-range
        // lanes, with their non-finite values at zero separation, add nothing
        if (in) {
            ++nIn;
            sT = fma(d, q.w, sT);
            sW += q.w;
            if (!MZ) sK += q.wk;
            if (KIND == SHEAR) {
                sG1 = fma(fma(q.s1w, u, q.s2w2 * v), inv, sG1);
                sG2 = fma(fma(q.s2w, u, -(s1w2_of<T, SPH>(q) * v)), inv, sG2);
            } else {
                sG1 += q.s1w;
                sG2 += q.s2w;
            }
            asm volatile("");
        }
    });
    const int lane = L.lane;
    constexpr int o = UPPER ? WQ : 0;
    wacc(w, o + QT, lane) += (double)sT;
    wacc(w, o + QG1, lane) += (double)sG1;
    wacc(w, o + QG2, lane) += (double)sG2;
    wacc(w, o + QW, lane) += (double)sW;
    if (!MZ) wacc(w, o + QK, lane) += (double)sK;
    tab_count(w, lane, UPPER ? w.wk + 1 : w.wk, n_bin, wave_sum_u32(L.valid ? nIn : 0u));
    if (UPPER) w.used1 = true;
    else w.used0 = true;
    w.edge = true;
}

// Multi pass: a segment whose pairs span more bins than an edge pass takes
// is swept once per S consecutive bins k0 .. k0 + S - 1.  Each pair's bin
// follows from the S + 1 edge tests of the sweep (FP32 d2 < lo; the lanes
// within the band of some edge re-decide every test in FP64, in a branch the
// wave takes only when such a lane exists), and the pair's terms are added to
// lane-private rows in LDS at that bin (the address carries the bin, so no
// VALU work per bin).  FP32: one float4 (Theta, G1, G2, W) per bin and lane,
// read, added to and written back (the LDS float atomics run at a small
// fraction of the rate of a b128 read and write), and the pair count with an
// integer LDS add; with K kept, a float2 (K, pair count) per bin and lane
// instead of the integer count.  FP64: rows[q][bin][lane] per quantity with
// FP64 LDS adds, and the integer counts.  The integer counts of bins 2 i and
// 2 i + 1 share a word (16 bits each: at most kSegMax < 2^16 pairs per lane
// and bin in a sweep).  The rows overlay the wave's window sums (a window is
// flushed before a multi segment), and a sweep's rows hold at most kSegMax
// terms per lane and bin, like the other passes' FP32 sums.  Bins per sweep:
// FP32 4 (3 with K kept), FP64 3 (2 with K kept), which fit the LDS the
// kernel's occupancy allows.
template <typename T, bool MZ>
constexpr int multi_bins() { return sizeof(T) == 4 ? (MZ ? 4 : 3) : (MZ ? 3 : 2); }

// the integer counts are packed two bins to a word (all but FP32 with K kept)
template <typename T, bool MZ>
constexpr bool multi_packed() { return !(sizeof(T) == 4 && !MZ); }

template <typename T, bool MZ>
constexpr int multi_row_bytes()
{
    constexpr int S = multi_bins<T, MZ>(), words = (S + 1) / 2;
    if constexpr (sizeof(T) == 4) return S * kWave * (MZ ? 16 : 24) + (MZ ? words * kWave * 4 : 0);
    else return S * kWave * win_q<MZ>() * 8 + words * kWave * 4;
}

// 8-byte units of a wave's LDS block: the FP64 window sums or the multi-pass rows
template <typename T, bool MZ>
constexpr int wave_lds_doubles()
{
    constexpr int win = (2 * win_q<MZ>() + 1) * kWave;
    constexpr int rows = (multi_row_bytes<T, MZ>() + 7) / 8;
    return win > rows ? win : rows;
}

// One sweep over a segment for bins k0 .. k0 + S - 1 (edges clamped at
// n_bin: a bin past the last one is empty).  ZCHECK: the z test,
// src/tpcf.f90:322-324.  ZSEP: shear with x_min == 0 and overlapping boxes,
// where a zero separation is a pair of bin 0: its position angle is 0/0, so
// its G terms are NaN (the reference's G_t, G_x of bin 0 and the lens' xi
// marks become NaN, src/tpcf.f90:343-355) and its separation is 0.
template <typename T, int KIND, bool MZ, bool ZCHECK, bool ZSEP, int SPH>
__device__ __forceinline__ void multi_sweep(Win &w, const LaneData<T> &L, const PairArgs<T> &A,
                                            const SrcRec<T> *__restrict__ r, const T *__restrict__ zr,
                                            uint32_t count, int n_bin, int k0)
{
    constexpr int S = multi_bins<T, MZ>(), QR = win_q<MZ>();
    const int lane = L.lane;
    // FP32 rows: float4 r4[S][kWave], then u32 cnt[(S + 1) / 2][kWave] holding
    // the counts of bins 2 i and 2 i + 1 in its low and high 16 bits (MZ) or
    // float2 r2[S][kWave] = (K, count); FP64 rows: T rows[QR][S][kWave], then
    // the packed counts
    float4 *const r4 = reinterpret_cast<float4 *>(w.acc);
    float2 *const r2 = reinterpret_cast<float2 *>(r4 + S * kWave);
    T *const rows = reinterpret_cast<T *>(w.acc);
    uint32_t *const cnt = sizeof(T) == 4 ? reinterpret_cast<uint32_t *>(r4 + S * kWave)
                                         : reinterpret_cast<uint32_t *>(rows + QR * S * kWave);
    if constexpr (sizeof(T) == 4) {
#pragma unroll
        for (int i = 0; i < S; ++i) {
            r4[i * kWave + lane] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (!MZ) r2[i * kWave + lane] = make_float2(0.f, 0.f);
            else if (i % 2 == 0) cnt[i / 2 * kWave + lane] = 0u;
        }
    } else {
#pragma unroll
        for (int i = 0; i < QR * S; ++i) rows[i * kWave + lane] = (T)0;
#pragma unroll
        for (int i = 0; i < S; i += 2) cnt[i / 2 * kWave + lane] = 0u;
    }
    // bin k counts in half k % 2 of word k / 2; the edge masks are nested
    // (ge[i + 1] within ge[i]), so k >= 2 is ge[2] and k is odd where the
    // exclusive or of ge[1] .. ge[S - 1] holds
    auto count_pair = [&](const uint64_t *ge) __attribute__((always_inline)) {
        static_assert(S >= 2 && S <= 4, "two count words");
        const bool up = S > 2 && __builtin_amdgcn_inverse_ballot_w64(ge[2 < S ? 2 : 1]);
        uint64_t par = 0;
#pragma unroll
        for (int i = 1; i < S; ++i) par ^= ge[i];
        const bool odd = __builtin_amdgcn_inverse_ballot_w64(par);
        atomicAdd(&cnt[(up ? kWave : 0) + lane], odd ? 0x10000u : 1u);
    };
    // the sweep's edges (FP32: the band [lo, hi] around each; FP64: the edge)
    T lo[S + 1], hi[S + 1];
#pragma unroll
    for (int j = 0; j <= S; ++j) {
        const int e = min(k0 + j, n_bin);
        if constexpr (sizeof(T) == 4) {
            lo[j] = A.e2f[e] - A.band * A.e2f[e];
            hi[j] = A.e2f[e] + A.band * A.e2f[e];
        } else {
            lo[j] = (T)A.e2d[e];
            hi[j] = lo[j];
        }
        // held in vector registers: the scalar registers are the kernel's
        // scarcer resource (spilled values would be reloaded in every pair loop)
        if constexpr (sizeof(T) == 4) asm volatile("" : "+v"(lo[j]), "+v"(hi[j]));
        else asm volatile("" : "+v"(lo[j]));
    }
    uint32_t j = 0;
    stream2<T>(r, count, [&](const SrcRec<T> &q) __attribute__((always_inline)) {
        const PairGeo<T, SPH> g = pair_geo<T, SPH>(L, q);
        const T d2 = g.d2;
        // lane masks (wave ballots of the compares, kept in scalar registers):
        // ge[i] = the lanes whose d2 is at or above edge k0 + i
        uint64_t ge[S + 1];
        if constexpr (sizeof(T) == 4) {
            uint64_t amb = 0;
#pragma unroll
            for (int i = 0; i <= S; ++i) {
                ge[i] = __builtin_amdgcn_ballot_w64(!(d2 < lo[i]));
                amb |= ge[i] & __builtin_amdgcn_ballot_w64(d2 <= hi[i]);
            }
            if (amb) {
                // rare (a wave-uniform branch): the band's lanes take every
                // test in FP64
                float xa = L.xa, ya = L.ya;
                asm volatile("" : "+v"(xa), "+v"(ya));
                const double dx = (double)xa - (double)q.x, dy = (double)ya - (double)q.y;
                const double dd2 = fma(dx, dx, dy * dy);
#pragma unroll
                for (int i = 0; i <= S; ++i)
                    ge[i] = (ge[i] & ~amb) | (__builtin_amdgcn_ballot_w64(!(dd2 < A.e2d[min(k0 + i, n_bin)])) & amb);
            }
        } else {
#pragma unroll
            for (int i = 0; i <= S; ++i) ge[i] = __builtin_amdgcn_ballot_w64(!(d2 < lo[i]));
        }
        uint64_t in = ge[0] & ~ge[S] & w.valid;
        if constexpr (KIND == SHEAR && ZCHECK) {
            typedef const __attribute__((address_space(4))) T CT;
            in &= __builtin_amdgcn_ballot_w64(!(L.za > ((CT *)zr)[j]));   // source in front of the lens
        }
        ++j;
        if (__builtin_amdgcn_inverse_ballot_w64(in)) {
            int k = 0;
#pragma unroll
            for (int i = 1; i < S; ++i) k += __builtin_amdgcn_inverse_ballot_w64(ge[i]) ? 1 : 0;
            // the pair's terms: t1 inv, t2 inv (shear: 1/d^2 applied by the
            // row update's FMA; kappa: inv = 1)
            T d, t1, t2, inv = (T)1;
            if constexpr (KIND == SHEAR) {
                T u, v;
                pair_proj<T, SPH, false>(L, g, d, inv, u, v);
                t1 = fma(q.s1w, u, q.s2w2 * v);
                t2 = fma(q.s2w, u, -(s1w2_of<T, SPH>(q) * v));
                if constexpr (ZSEP) {
                    const bool z0 = d2 == (T)0;
                    d = z0 ? (T)0 : d;
                    t1 = z0 ? (T)__builtin_nan("") : t1;
                    t2 = z0 ? (T)__builtin_nan("") : t2;
                }
            } else {
                if constexpr (SPH) d = sph_angle<sph_short(SPH)>(d2);
                else d = Geo<T>::dist(d2);
                t1 = q.s1w;
                t2 = q.s2w;
            }
            const int at = k * kWave + lane;
            if constexpr (sizeof(T) == 4) {
                float4 a = r4[at];
                a.x = fma(d, q.w, a.x);
                if constexpr (KIND == SHEAR) {
                    a.y = fma(t1, inv, a.y);
                    a.z = fma(t2, inv, a.z);
                } else {
                    a.y += t1;
                    a.z += t2;
                }
                a.w += q.w;
                r4[at] = a;
                if constexpr (MZ) {
                    count_pair(ge);
                } else {
                    float2 b = r2[at];
                    b.x += q.wk;
                    b.y += 1.0f;   // exact: at most kSegMax per sweep
                    r2[at] = b;
                }
            } else {
                T *const p = rows + at;
                atomicAdd(p + QT * S * kWave, d * q.w);
                atomicAdd(p + QG1 * S * kWave, KIND == SHEAR ? t1 * inv : t1);
                atomicAdd(p + QG2 * S * kWave, KIND == SHEAR ? t2 * inv : t2);
                atomicAdd(p + QW * S * kWave, q.w);
                if constexpr (!MZ) atomicAdd(p + QK * S * kWave, q.wk);
                count_pair(ge);
            }
        }
    });
}

// The sweep's rows into the per-bin registers (lane b holds bin b) and, with
// marks, into the marks of the chunk's lenses, as win_flush does for a window.
template <typename T, int KIND, bool MARKS, bool MZ>
__device__ __forceinline__ void multi_flush(Win &w, const LaneData<T> &L, int n_bin, double *__restrict__ marks,
                                            int k0)
{
    constexpr int S = multi_bins<T, MZ>(), QR = win_q<MZ>();
    const float4 *r4 = reinterpret_cast<const float4 *>(w.acc);
    const float2 *r2 = reinterpret_cast<const float2 *>(r4 + S * kWave);
    const T *rows = reinterpret_cast<const T *>(w.acc);
    const uint32_t *cnt = sizeof(T) == 4 ? reinterpret_cast<const uint32_t *>(r4 + S * kWave)
                                         : reinterpret_cast<const uint32_t *>(rows + QR * S * kWave);
    const int lane = L.lane;
    const double dwa = (double)L.wa, daa = (double)L.aa;
    const double gs = (KIND == SHEAR) ? -1.0 : 1.0;   // G_t -= ... for shear, src/tpcf.f90:348-349
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int bin = k0 + s;
        if (bin >= n_bin) break;
        const int at = s * kWave + lane;
        double aT, aG1, aG2, aW, aK;
        uint32_t nl;
        if constexpr (sizeof(T) == 4) {
            const float4 a = r4[at];
            aT = a.x;
            aG1 = a.y;
            aG2 = a.z;
            aW = a.w;
            if constexpr (MZ) {
                nl = (cnt[s / 2 * kWave + lane] >> (16 * (s % 2))) & 0xffffu;
                aK = aW;
            } else {
                const float2 b = r2[at];
                aK = b.x;
                nl = (uint32_t)b.y;
            }
        } else {
            aT = rows[QT * S * kWave + at];
            aG1 = rows[QG1 * S * kWave + at];
            aG2 = rows[QG2 * S * kWave + at];
            aW = rows[QW * S * kWave + at];
            aK = MZ ? aW : (double)rows[QK * S * kWave + at];
            nl = (cnt[s / 2 * kWave + lane] >> (16 * (s % 2))) & 0xffffu;
        }
        if (!__ballot(nl != 0u)) continue;   // no pair of the chunk in this bin
        const uint32_t n = wave_sum_u32(nl);
        const double sT = wave_sum(dwa * aT);
        const double sG1 = gs * wave_sum(daa * aG1);
        const double sG2 = gs * wave_sum(daa * aG2);
        const double sW = wave_sum(dwa * aW);
        const double sK = MZ ? sW : wave_sum(dwa * aK);
        if (lane == bin) {
            w.t[QT] += sT;
            w.t[QG1] += sG1;
            w.t[QG2] += sG2;
            w.t[QW] += sW;
            if (!MZ) w.t[QK] += sK;
            w.tn += n;
        }
        if (MARKS && L.valid) {
            const double m0 = gs * daa * aG1, m1 = gs * daa * aG2, m2 = dwa * aK, m3 = dwa * aW;
            if (m0 != 0.0 || m1 != 0.0 || m2 != 0.0 || m3 != 0.0) {   // NaN != 0
                const int64_t nm = w.mstride;
                double *mk = marks + mark_at(nm, bin, 0, L.li);
                unsafeAtomicAdd(&mk[0], m0);
                unsafeAtomicAdd(&mk[nm], m1);
                if (!MZ) unsafeAtomicAdd(&mk[2 * nm], m2);
                unsafeAtomicAdd(&mk[3 * nm], m3);
            }
        }
    }
}



// The work of one item: lens chunk item.chunk against the source segments in
// its band of cell rows, into the per-XCD accumulator copy `slot`.
// Runs f with the spherical variant of a segment visit: SPH + 4 (the short
// angle series) when every chord of the visit lies below kSphShort (hi2 is
// an upper bound on the visit's squared chords), else SPH.
template <int SPH, typename T, typename F>
__device__ __forceinline__ void sph_visit(T hi2, F &&f)
{
    static_assert(SPH == 1 || SPH == 2, "spherical passes only");
    if (hi2 < (T)kSphShort) f(std::integral_constant<int, SPH | 4>());
    else f(std::integral_constant<int, SPH>());
}

// Fast class: every pair of the segment is in one bin, so only the terms
// that depend on the pair are summed here (Theta, and G1, G2 for shear); the
// others come from the segment's sums.
template <typename T, int KIND, int SPH>
__device__ __forceinline__ void fast_sums(const LaneData<T> &L, const SrcRec<T> *__restrict__ r, uint32_t cnt, T &sT,
                                          T &sG1, T &sG2)
{
    if constexpr (KIND == SHEAR) {
        stream2<T>(r, cnt, [&](const SrcRec<T> &q) __attribute__((always_inline)) {
            const PairGeo<T, SPH> g = pair_geo<T, SPH>(L, q);
            T d, inv, u, v;
            pair_proj<T, SPH, false>(L, g, d, inv, u, v);
            sT = fma(d, q.w, sT);
            sG1 = fma(fma(q.s1w, u, q.s2w2 * v), inv, sG1);
            sG2 = fma(fma(q.s2w, u, -(s1w2_of<T, SPH>(q) * v)), inv, sG2);
        });
    } else {
        stream2<T>(r, cnt, [&](const SrcRec<T> &q) __attribute__((always_inline)) {
            const PairGeo<T, SPH> g = pair_geo<T, SPH>(L, q);
            if constexpr (SPH) sT = fma(sph_angle<sph_short(SPH)>(g.d2), q.w, sT);
            else sT = fma(Geo<T>::dist(g.d2), q.w, sT);
        });
    }
}

template <typename T, int KIND, bool MARKS, bool MZ, int SPH>
__device__ __forceinline__ void pair_item(const PairArgs<T> &A, const CatArgs<T> &C, const Item item, const int slot,
                                          const int lane, double *acc, const T e2_lane,
                                          unsigned long long *__restrict__ probe)
{
    constexpr int WQ = win_q<MZ>();
    const Chunk<T> ch = C.chunks[item.chunk];
    const int row0 = item.row0, row1 = item.row1;   // this item's band of cell rows
    LaneData<T> L;
    L.lane = lane;
    L.valid = lane < (int)ch.count;
    L.li = ch.start + (L.valid ? lane : 0);
    L.xa = C.lx[L.li];
    L.ya = C.ly[L.li];
    L.za = (KIND == SHEAR) ? C.lz[L.li] : (T)0;
    L.wa = L.valid ? C.lw[L.li] : (T)0;
    L.aa = L.valid ? C.la[L.li] : (T)0;
    L.vmask = L.valid ? (T)1 : (T)0;
    double2 chp = make_double2(0.0, 0.0);
    if constexpr (SPH) {
        const double *f = C.sph_lf + 6 * (size_t)L.li;   // [n_lens][east 3, north 3]
        L.pa = C.sph_lp[L.li];
        L.fe0 = f[0]; L.fe1 = f[1]; L.fe2 = f[2];
        L.fn0 = f[3]; L.fn1 = f[4]; L.fn2 = f[5];
        chp = C.sph_chp[item.chunk];
    }
    const int n_bin = A.n_bin;

    Win w;
    w.acc = acc;
    win_clear<MZ>(w, lane);
    w.pb = -1;
    w.wk = -3;
    w.valid = __ballot(L.valid);
    w.n_valid = ch.count;
    w.mstride = C.marks_n;
    for (int q = 0; q < kNQ; ++q) w.t[q] = 0.0;
    w.tn = 0;

    const double xm = A.x_max, mg = C.margin;
    // The grid and the catalogue's segment pointers are kept in this wave's
    // LDS block (after its sums) and read again where the row and segment
    // loops use them: held in scalar registers across the item they displaced
    // values that the pair loops then reloaded on every iteration.  The
    // pointers are read back as constant-address-space pointers (the records
    // they address are not written while the kernel runs), so their loads
    // stay scalar.
    typedef __attribute__((address_space(3))) double LdsD;
    LdsD *const gsh = (LdsD *)(acc + wave_lds_doubles<T, MZ>());
    if (lane == 0) {
        gsh[0] = C.g.x0;
        gsh[1] = C.g.y0;
        gsh[2] = C.g.cell;
        gsh[3] = C.g.inv;
        gsh[4] = __longlong_as_double(((long long)C.g.ny << 32) | (unsigned)C.g.nx);
        gsh[5] = __longlong_as_double((long long)C.g.ncells);
        gsh[6] = __longlong_as_double((long long)(uintptr_t)C.segs);
        gsh[7] = __longlong_as_double((long long)(uintptr_t)C.seg_first);
        gsh[8] = __longlong_as_double((long long)(uintptr_t)C.srec);
        gsh[9] = __longlong_as_double((long long)(uintptr_t)C.sz);
        gsh[10] = __longlong_as_double((long long)(uintptr_t)C.pbox);
    }
    typedef const __attribute__((address_space(4))) Seg<T> CSeg;
    typedef const __attribute__((address_space(4))) uint32_t CU32;
    typedef const __attribute__((address_space(4))) int4 CI4;
    auto lds_ptr = [&](int k) __attribute__((always_inline)) {
        LdsD *q = gsh;
        asm volatile("" : "+s"(q));   // read where used, not hoisted
        return (uintptr_t)__double_as_longlong(q[k]);
    };
    for (int pb = 0; pb < C.n_pb; ++pb) {
        CI4 *const pbq = (CI4 *)lds_ptr(10) + pb;
        const int4 pbx = make_int4(pbq->x, pbq->y, pbq->z, pbq->w);
        if (pbx.x > pbx.y) continue;   // no sources in this patch
        const int cy0 = max(row0, pbx.z);
        const int cy1 = min(row1, pbx.w);
        if (cy0 > cy1) continue;
        for (int cy = cy0; cy <= cy1; ++cy) {
            int cx0, cx1;
            LdsD *gq = gsh;
            asm volatile("" : "+s"(gq));   // read at each row, not hoisted
            Grid g;
            g.x0 = gq[0];
            g.y0 = gq[1];
            g.cell = gq[2];
            g.inv = gq[3];
            const long long nxy = __double_as_longlong(gq[4]);
            g.nx = (int32_t)(nxy & 0xffffffffll);
            g.ny = (int32_t)(nxy >> 32);
            g.ncells = (int64_t)__double_as_longlong(gq[5]);
            // opaque copies: the FP64 box is formed per row instead of being
            // hoisted into eight registers live across the whole loop nest
            T bx0 = ch.bx0, bx1 = ch.bx1, by0 = ch.by0, by1 = ch.by1;
            asm volatile("" : "+s"(bx0), "+s"(bx1), "+s"(by0), "+s"(by1));
            if (!col_span(g, (double)bx0, (double)bx1, (double)by0, (double)by1, cy, xm, mg, cx0, cx1)) continue;
            cx0 = max(cx0, pbx.x);
            cx1 = min(cx1, pbx.y);
            if (cx0 > cx1) continue;
            const int64_t key0 = (int64_t)pb * g.ncells + (int64_t)cy * g.nx + cx0;
            CU32 *const seg_first = (CU32 *)lds_ptr(7);
            const uint32_t s_begin = __builtin_amdgcn_readfirstlane(seg_first[key0]);
            const uint32_t s_end = __builtin_amdgcn_readfirstlane(seg_first[key0 + (cx1 - cx0) + 1]);
            for (uint32_t si = s_begin; si < s_end; ++si) {
                CSeg *const sq = (CSeg *)lds_ptr(6) + si;
                Seg<T> sg;
                sg.start = sq->start; sg.count = sq->count;
                sg.x0 = sq->x0; sg.x1 = sq->x1; sg.y0 = sq->y0; sg.y1 = sq->y1; sg.zmin = sq->zmin;
                sg.W = sq->W; sg.K = sq->K; sg.S1 = sq->S1; sg.S2 = sq->S2;
                const uint32_t cnt = __builtin_amdgcn_readfirstlane(sg.count);
                const SrcRec<T> *__restrict__ r =
                    (const SrcRec<T> *)lds_ptr(8) + __builtin_amdgcn_readfirstlane(sg.start);
                // Warm L2 with the segment's records: lane l reads one value of
                // each 128-B line it covers (FP32, 32-B records: record 4 l, so
                // 64 lanes cover 256 records, a whole segment; FP64, 64-B
                // records: records 2 l and 2 l + 128; segments are never
                // empty).  The scalar loads of the passes wait for every load
                // in flight, so their latency is exposed once per two records;
                // with the lines on their way to L2 it is the L2 latency
                // rather than HBM's (2 % on the 20M x 20M workload).  The
                // values are only consumed after the pass.
                // (global-address-space loads: a flat load would also hold up
                // the waits of the scalar loads)
                typedef const __attribute__((address_space(1))) float GF;
                constexpr uint32_t kRecLine = 128 / sizeof(SrcRec<T>);
                const uint32_t pfi = min(kRecLine * (uint32_t)lane, cnt - 1);
                const float pfv = ((GF *)(r + pfi))[0];
                float pfv2 = 0.0f;
                if constexpr (kRecLine * kWave < kSegMax) {
                    const uint32_t pfj = min(kRecLine * (uint32_t)lane + kRecLine * kWave, cnt - 1);
                    pfv2 = ((GF *)(r + pfj))[0];
                }
                // squared distance range between the chunk box and the segment box
                const T dxl = fmax((T)0, fmax(sg.x0 - ch.bx1, ch.bx0 - sg.x1));
                const T dyl2 = fmax((T)0, fmax(sg.y0 - ch.by1, ch.by0 - sg.y1));
                const T dxh = fmax(sg.x1 - ch.bx0, ch.bx1 - sg.x0);
                const T dyh = fmax(sg.y1 - ch.by0, ch.by1 - sg.y0);
                T lo2 = dxl * dxl + dyl2 * dyl2;
                T hi2 = dxh * dxh + dyh * dyh;
                if constexpr (SPH) {
                    // the third component: chord bounds between the 3-D boxes
                    const double2 sp = C.sph_segp[si];
                    const T dpl = fmax((T)0, fmax(sp.x - chp.y, chp.x - sp.y));
                    const T dph = fmax(sp.y - chp.x, chp.y - sp.x);
                    lo2 = fma(dpl, dpl, lo2);
                    hi2 = fma(dph, dph, hi2);
                }
                lo2 *= A.cls_lo;
                hi2 *= A.cls_hi;
                const int kmin = __popcll(__ballot(e2_lane <= lo2)) - 1;
                const int kmax = __popcll(__ballot(e2_lane <= hi2)) - 1;
                if (kmax < 0 || kmin >= n_bin) continue;   // all pairs outside [x_min, x_max]
                const bool zc = (KIND == SHEAR) && (ch.zmax > sg.zmin);
                int cls;
                if (kmin == kmax && !zc && lo2 > (T)0) {
                    cls = 0;
                    // ---- fast: the whole segment is in bin kmin ----
                    if (!(w.open && pb == w.pb && (kmin == w.wk || kmin == w.wk + 1))) {
                        // a walk moving to lower bins keeps the bin below in
                        // the new window, one moving up the bin above
                        const int k0 = (w.open && pb == w.pb && kmin < w.wk) ? kmin - 1 : kmin;
                        win_open<T, KIND, MARKS, MZ>(w, L, n_bin, C.marks, pb, k0);
                    }
                    T sT = 0, sG1 = 0, sG2 = 0;
                    if constexpr (SPH == 0) {
                        if (KIND == SHEAR) {
                            stream2<T>(r, cnt, [&](const SrcRec<T> &q) __attribute__((always_inline)) {
                                const PairGeo<T, SPH> g = pair_geo<T, SPH>(L, q);
                                T d, inv, u, v;
                                pair_proj<T, SPH, false>(L, g, d, inv, u, v);
                                sT = fma(d, q.w, sT);
                                sG1 = fma(fma(q.s1w, u, q.s2w2 * v), inv, sG1);
                                sG2 = fma(fma(q.s2w, u, -(s1w2_of<T, SPH>(q) * v)), inv, sG2);
                            });
                        } else {
                            stream2<T>(r, cnt, [&](const SrcRec<T> &q) __attribute__((always_inline)) {
                                const PairGeo<T, SPH> g = pair_geo<T, SPH>(L, q);
                                sT = fma(Geo<T>::dist(g.d2), q.w, sT);
                            });
                        }
                    } else {
                        sph_visit<SPH>(hi2, [&](auto S) __attribute__((always_inline)) {
                            fast_sums<T, KIND, decltype(S)::value>(L, r, cnt, sT, sG1, sG2);
                        });
                    }
                    const int o = (kmin - w.wk) * WQ;
                    if (o == 0) w.used0 = true; else w.used1 = true;
                    wacc(w, o + QT, lane) += (double)sT;
                    if (KIND == SHEAR) {
                        wacc(w, o + QG1, lane) += (double)sG1;
                        wacc(w, o + QG2, lane) += (double)sG2;
                    }
                    if (MARKS) {
                        wacc(w, o + QW, lane) += sg.W;
                        if (!MZ) wacc(w, o + QK, lane) += sg.K;
                        if (KIND == KAPPA) {
                            wacc(w, o + QG1, lane) += sg.S1;
                            wacc(w, o + QG2, lane) += sg.S2;
                        }
                    }
                    // pair count (and, without marks, W, K and kappa's G) from the segment sums
                    if (lane == kmin) {
                        w.tn += (uint64_t)cnt * ch.count;
                        if (!MARKS) {
                            w.t[QW] += ch.wsum * sg.W;
                            if (!MZ) w.t[QK] += ch.wsum * sg.K;
                            if (KIND == KAPPA) {
                                w.t[QG1] += ch.asum * sg.S1;
                                w.t[QG2] += ch.asum * sg.S2;
                            }
                        }
                    }
                } else if (kmax == kmin + 1 && !zc) {
                    cls = 1;
                    // ---- edge: one bin edge inside the distance range ----
                    win_open<T, KIND, MARKS, MZ>(w, L, n_bin, C.marks, pb, kmin);
                    if (kmax >= n_bin && kmin >= 0 && !(A.xmin_zero && lo2 == (T)0)) {
                        // x_max inside the range: only bin n_bin - 1 counts (the
                        // edge pass would derive it as total - the pairs beyond
                        // x_max; at x_min it sums bin 0 directly already)
                        if constexpr (SPH == 0) {
                            edge_side_pass<T, KIND, MZ, false, SPH>(w, L, A, r, cnt, n_bin);
                        } else {
                            sph_visit<SPH>(hi2, [&](auto S) __attribute__((always_inline)) {
                                edge_side_pass<T, KIND, MZ, false, decltype(S)::value>(w, L, A, r, cnt, n_bin);
                            });
                        }
                    } else if (lo2 > (T)0) {
                        if constexpr (SPH == 0) {
                            edge_pass<T, KIND, MZ, false, false, SPH>(w, L, A, r, cnt, n_bin);
                        } else {
                            sph_visit<SPH>(hi2, [&](auto S) __attribute__((always_inline)) {
                                edge_pass<T, KIND, MZ, false, false, decltype(S)::value>(w, L, A, r, cnt, n_bin);
                            });
                        }
                    } else if (KIND == SHEAR && A.xmin_zero && kmin == 0) {
                        if constexpr (SPH == 0) {
                            edge_pass<T, KIND, MZ, true, true, SPH>(w, L, A, r, cnt, n_bin);
                        } else {
                            sph_visit<SPH>(hi2, [&](auto S) __attribute__((always_inline)) {
                                edge_pass<T, KIND, MZ, true, true, decltype(S)::value>(w, L, A, r, cnt, n_bin);
                            });
                        }
                    } else {
                        if constexpr (SPH == 0) {
                            edge_pass<T, KIND, MZ, true, false, SPH>(w, L, A, r, cnt, n_bin);
                        } else {
                            sph_visit<SPH>(hi2, [&](auto S) __attribute__((always_inline)) {
                                edge_pass<T, KIND, MZ, true, false, decltype(S)::value>(w, L, A, r, cnt, n_bin);
                            });
                        }
                    }
                } else {
                    cls = 2;
                    // ---- multi: sweeps of multi_bins() bins over [kmin, kmax] ----
                    const T *zr = (KIND == SHEAR) ? (const T *)lds_ptr(9) + __builtin_amdgcn_readfirstlane(sg.start) : nullptr;
                    // the sweeps' rows overlay the window sums
                    win_flush<T, KIND, MARKS, MZ>(w, L, n_bin, C.marks);
                    const int wlast = min(kmax, n_bin - 1);
                    for (int k0 = max(kmin, 0); k0 <= wlast; k0 += multi_bins<T, MZ>()) {
                        // a zero separation is in range only at x_min == 0, in
                        // bin 0, with the boxes overlapping
                        const bool zsep = KIND == SHEAR && A.xmin_zero && k0 == 0 && lo2 == (T)0;
                        if constexpr (SPH == 0) {
                            if (zc) {
                                if (zsep) multi_sweep<T, KIND, MZ, true, true, SPH>(w, L, A, r, zr, cnt, n_bin, k0);
                                else multi_sweep<T, KIND, MZ, true, false, SPH>(w, L, A, r, zr, cnt, n_bin, k0);
                            } else {
                                if (zsep) multi_sweep<T, KIND, MZ, false, true, SPH>(w, L, A, r, zr, cnt, n_bin, k0);
                                else multi_sweep<T, KIND, MZ, false, false, SPH>(w, L, A, r, zr, cnt, n_bin, k0);
                            }
                        } else {
                            sph_visit<SPH>(hi2, [&](auto S) __attribute__((always_inline)) {
                                constexpr int S2 = decltype(S)::value;
                                if (zc) {
                                    if (zsep) multi_sweep<T, KIND, MZ, true, true, S2>(w, L, A, r, zr, cnt, n_bin, k0);
                                    else multi_sweep<T, KIND, MZ, true, false, S2>(w, L, A, r, zr, cnt, n_bin, k0);
                                } else {
                                    if (zsep) multi_sweep<T, KIND, MZ, false, true, S2>(w, L, A, r, zr, cnt, n_bin, k0);
                                    else multi_sweep<T, KIND, MZ, false, false, S2>(w, L, A, r, zr, cnt, n_bin, k0);
                                }
                            });
                        }
                        multi_flush<T, KIND, MARKS, MZ>(w, L, n_bin, C.marks, k0);
                    }
                    win_clear<MZ>(w, lane);   // the window sums start from zero again
                }
                asm volatile("" ::"v"(pfv), "v"(pfv2));
                if (A.counters) {
                    // class counters (diagnostic runs only): added per segment,
                    // so that no register is held for them across the item
                    if (lane == 0) {
                        atomicAdd(&probe[cls], 1ull);
                        atomicAdd(&probe[3 + cls], (unsigned long long)cnt * ch.count);
                    }
                }
            }
        }
        win_flush<T, KIND, MARKS, MZ>(w, L, n_bin, C.marks);
        // hand the patch's per-bin sums to the patch-pair accumulators (rows
        // of all the call's bins; pp and ppn point at this pass's first bin)
        if (lane < n_bin && w.tn > 0) {
            const size_t o = (((size_t)slot * C.n_pa + ch.pa) * C.n_pb + pb) * A.pp_stride + lane;
            unsafeAtomicAdd(&C.pp[o * kNQ + 0], w.t[QT]);
            unsafeAtomicAdd(&C.pp[o * kNQ + 1], w.t[QG1]);
            unsafeAtomicAdd(&C.pp[o * kNQ + 2], w.t[QG2]);
            unsafeAtomicAdd(&C.pp[o * kNQ + 3], w.t[QW]);
            unsafeAtomicAdd(&C.pp[o * kNQ + 4], MZ ? w.t[QW] : w.t[QK]);
            atomicAdd(&C.ppn[o], (unsigned long long)w.tn);
        }
        for (int q = 0; q < kNQ; ++q) w.t[q] = 0.0;
        w.tn = 0;
    }
}

// Waves per SIMD the register budget is set for.  FP32: 7 waves (72 VGPRs,
// 94 SGPRs; the LDS of 4-wave workgroups, 18.2 KB with all m == 0 and 22.7 KB
// otherwise, allows 8 and 7 of them per CU).  The kernel issues VALU work
// back to back, and at 8 waves (64 VGPRs, 78 SGPRs) its per-visit code
// spills registers, to vector lanes and to scratch, in instructions that
// take VALU issue slots: on the 20M x 20M shear workload the pair kernel
// took 3512 ms at 6 waves, 3506 ms at 7 and 3536 ms at 8 (3537 and 3609 ms
// with marks at 7 and 8; the mock batch and the adjoint within 0.3 %), all
// on one GPU.  FP64 with all m == 0 runs 6 waves (80 VGPRs, some spills): on
// the FP64 1M x 2M shear workload 4 waves took 208 ms, 5 took 187 ms and 6
// took 182 ms (7: the same as 6, with twice the spills).  FP64 with some
// m != 0 runs 5 waves (96 VGPRs): on that workload with m drawn from
// N(0, 0.05), 203 ms at 4 waves, 185 ms at 5.
template <typename T, bool MZ>
constexpr int pair_waves_per_simd() { return sizeof(T) == 4 ? 7 : (MZ ? 6 : 5); }
// the spherical geometry (FP64) holds seven more lane values: 4 waves, up to
// 128 VGPRs (3 waves were slower on the 20M x 20M sphere workload)
template <typename T, bool MZ, int SPH>
constexpr int pair_waves() { return SPH ? 4 : pair_waves_per_simd<T, MZ>(); }

// One catalogue: one item per wave; workgroup b runs on XCD b % kSlots (the
// hardware deals workgroups to the XCDs round-robin) and takes the next items
// of that XCD's range, which the preparation ordered band by band, so the
// waves running together on an XCD read the same source rows.  The catalogue's
// pointers arrive as separate arguments (scalar loads on demand).
template <typename T, int KIND, bool MARKS, bool MZ, int SPH>
__global__ __launch_bounds__(kWave *kWavesPerBlock)
__attribute__((amdgpu_waves_per_eu(pair_waves<T, MZ, SPH>(), 8))) void k_pairs(
    const PairArgs<T> A, const Grid g, const int32_t n_pa, const int32_t n_pb, const double margin,
    const Chunk<T> *__restrict__ chunks, const T *__restrict__ lx, const T *__restrict__ ly,
    const T *__restrict__ lz, const T *__restrict__ la, const T *__restrict__ lw, const Seg<T> *__restrict__ segs,
    const uint32_t *__restrict__ seg_first, const int4 *__restrict__ pbox, const SrcRec<T> *__restrict__ srec,
    const T *__restrict__ sz, const T *__restrict__ e2lane, const Item *__restrict__ items, double *__restrict__ pp,
    unsigned long long *__restrict__ ppn, double *__restrict__ marks, const int64_t marks_n,
    unsigned long long *__restrict__ probe, const double *__restrict__ sph_lp, const double *__restrict__ sph_lf, const double2 *__restrict__ sph_segp,
    const double2 *__restrict__ sph_chp)
{
    static_assert(!SPH || sizeof(T) == 8, "spherical geometry runs in FP64");
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int64_t b = blockIdx.x;
    const int slot = (int)(b % kSlots);
    const int64_t it = ((b % kSlots) * A.per_xcd + b / kSlots) * kWavesPerBlock + wv;
    if (it >= A.n_items) return;
    // per wave: the window sums or multi-pass rows, then the grid and the
    // segment pointers (11 x 8 bytes)
    static_assert((wave_lds_doubles<T, MZ>() + 12) % 2 == 0, "16-byte aligned wave blocks (double2 reads)");
    __shared__ __align__(16) double s_acc[kWavesPerBlock][wave_lds_doubles<T, MZ>() + 12];
    CatArgs<T> C;
    C.g = g;
    C.n_pa = n_pa;
    C.n_pb = n_pb;
    C.margin = margin;
    C.chunks = chunks;
    C.lx = lx; C.ly = ly; C.lz = lz; C.la = la; C.lw = lw;
    C.segs = segs;
    C.seg_first = seg_first;
    C.pbox = pbox;
    C.srec = srec;
    C.sz = sz;
    C.pp = pp;
    C.ppn = ppn;
    C.marks = marks;
    C.marks_n = marks_n;
    C.sph_lp = sph_lp;
    C.sph_lf = sph_lf;
    C.sph_segp = sph_segp;
    C.sph_chp = sph_chp;
    const Item item = items[it];
    if (item.chunk == kNoChunk) return;   // padding of this XCD's item range
    pair_item<T, KIND, MARKS, MZ, SPH>(A, C, item, slot, lane, s_acc[wv], e2lane[lane], probe);
}

// The pair kernel's arguments for bin group `group` of bins: bins
// [g0, g1) with g0 = group kGroupBins, at most kGroupBins of them.  The
// group's edges are the call's edges g0..g1; d == x_max belongs to the last
// bin of the last group, and a pair at an inner group's top edge to the next
// group (both groups test that edge alike).  The squared edges one lane each
// go to slot `group` of the context's edge table (sized for every group of
// the call, so no later group reallocates it while a pass may run).
template <typename T>
int make_pair_args(tpcf_ctx *ctx, const tpcf_bins *bins, int group, PairArgs<T> &A, const T **e2lane_out)
{
    std::memset(&A, 0, sizeof A);
    const int n_all = bins->n_bin, n_groups = bin_groups(n_all);
    const int g0 = group * kGroupBins, g1 = std::min(n_all, g0 + kGroupBins);
    if (group < 0 || g0 >= n_all) return fail(ctx, TPCF_ERR_ARG, "bin group out of range");
    A.n_bin = g1 - g0;
    A.pp_stride = n_all;
    A.counters = ctx->opt_counters ? 1 : 0;
    const int n = A.n_bin;
    const double *edges = bins->edges + g0;
    A.x_max = edges[n];
    A.xmin_zero = (g0 == 0 && edges[0] == 0.0) ? 1 : 0;
    if (sizeof(T) == 4) {
        A.cls_lo = (T)(1.0 - 2e-6);
        A.cls_hi = (T)(1.0 + 2e-6);
    } else {
        A.cls_lo = (T)(1.0 - 1e-12);
        A.cls_hi = (T)(1.0 + 1e-12);
    }
    A.band = 1e-5f;
    for (int i = 0; i <= kGroupBins; ++i) {
        if (i <= n) {
            const double e = edges[i];
            A.e2d[i] = e * e;
            A.e2f[i] = (float)(e * e);
        } else {
            A.e2d[i] = INFINITY;
            A.e2f[i] = INFINITY;
        }
    }
    if (g1 == n_all) A.e2d[n] = std::nextafter(A.e2d[n], INFINITY);   // d == x_max belongs to the last bin
    T *e2lane;
    TPCF_TRY(ensure(ctx, tpcf_ctx::B_E2, sizeof(T) * kWave * n_groups, (void **)&e2lane));
    e2lane += (size_t)group * kWave;
    T h[kWave];
    for (int i = 0; i < kWave; ++i) h[i] = (i <= n) ? (sizeof(T) == 4 ? (T)A.e2f[i] : (T)A.e2d[i]) : (T)INFINITY;
    TPCF_HIP(ctx, hipMemcpyAsync(e2lane, h, sizeof h, hipMemcpyHostToDevice, ctx->stream));
    TPCF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *e2lane_out = e2lane;
    return TPCF_OK;
}

template <typename T>
CatArgs<T> cat_args(const Prepared<T> &p, int32_t n_pa, int32_t n_pb, double *pp, unsigned long long *ppn,
                    double *marks)
{
    CatArgs<T> C;
    std::memset(&C, 0, sizeof C);
    C.g = p.grid;
    C.n_pa = n_pa;
    C.n_pb = n_pb;
    C.margin = p.margin;
    C.chunks = p.chunks;
    C.lx = p.lx; C.ly = p.ly; C.lz = p.lz; C.la = p.la; C.lw = p.lw;
    C.segs = p.segs;
    C.seg_first = p.seg_first;
    C.pbox = p.pbox;
    C.srec = p.srec;
    C.sz = p.sz;
    C.pp = pp;
    C.ppn = ppn;
    C.marks = marks;
    C.marks_n = p.n_lens;
    C.sph_lp = p.sph_lp;
    C.sph_lf = p.sph_lf;
    C.sph_segp = p.sph_segp;
    C.sph_chp = p.sph_chp;
    C.sph_rev = p.sph_rev;
    return C;
}

template <typename T>
int run_pairs(tpcf_ctx *ctx, const CallArgs<T> &a, const Prepared<T> &p, double *pp, unsigned long long *ppn,
              double *marks)
{
    const tpcf_bins *kb = a.kbins ? a.kbins : a.bins;
    unsigned long long *probe = nullptr;
    if (ctx->opt_counters) {
        TPCF_TRY(ensure(ctx, tpcf_ctx::B_PROBE, sizeof(unsigned long long) * 8, (void **)&probe));
        fill_u64(ctx->stream, probe, 8);
    }
    if (p.n_items == 0) return TPCF_OK;
    const int64_t nblk = (p.n_items + kWavesPerBlock - 1) / kWavesPerBlock;
    // one pass per group of up to kGroupBins bins, each into its bins' columns
    // of the accumulators and marks planes
    for (int g = 0; g < bin_groups(kb->n_bin); ++g) {
        PairArgs<T> A;
        const T *e2lane;
        TPCF_TRY(make_pair_args<T>(ctx, kb, g, A, &e2lane));
        A.n_items = p.n_items;
        A.per_xcd = (nblk + kSlots - 1) / kSlots;
        const size_t g0 = (size_t)g * kGroupBins;
        const CatArgs<T> C = cat_args<T>(p, a.n_pa, a.n_pb, pp + g0 * kNQ, ppn + g0,
                                         marks ? marks + g0 * 4 * (size_t)p.n_lens : nullptr);
        if (g == 0) TPCF_HIP(ctx, hipEventRecord(ctx->ev[2], ctx->stream));
        TPCF_TRY(launch_pairs<T>(ctx, ctx->stream, a.kind, p.m_zero, a.marks, A, C, e2lane, p.items, probe));
    }
    TPCF_HIP(ctx, hipEventRecord(ctx->ev[3], ctx->stream));
    return TPCF_OK;
}

// The pair kernel for one catalogue's items: only a kernel launch on `s`
// (graph-capturable).  A.n_items and A.per_xcd describe the items.
template <typename T>
int launch_pairs(tpcf_ctx *ctx, hipStream_t s, int kind, bool m_zero, bool marks, const PairArgs<T> &A,
                 const CatArgs<T> &C, const T *e2lane, const Item *items, unsigned long long *probe)
{
    if (A.n_items == 0) return TPCF_OK;
    const dim3 grid((unsigned)(A.per_xcd * kSlots)), block(kWave * kWavesPerBlock);
#define TPCF_LAUNCH(K, M, Z, S)                                                                                   \
    hipLaunchKernelGGL((k_pairs<T, K, M, Z, S>), grid, block, 0, s, A, C.g, C.n_pa, C.n_pb, C.margin, C.chunks,     \
                       C.lx, C.ly, C.lz, C.la, C.lw, C.segs, C.seg_first, C.pbox, C.srec, C.sz, e2lane, items, C.pp, \
                       C.ppn, C.marks, C.marks_n, probe, C.sph_lp, C.sph_lf, C.sph_segp, C.sph_chp)
    if (C.sph_lp) {
        // spherical geometry: FP64 only
        if constexpr (sizeof(T) == 8) {
            if (C.sph_rev && kind == SHEAR) {
                // the adjoint's source-side pass: marks, unit m == 0 stream
                if (!(marks && m_zero))
                    return fail(ctx, TPCF_ERR_UNSUPPORTED, "a swapped spherical pass computes marks with m == 0");
                TPCF_LAUNCH(SHEAR, true, true, 2);
            } else if (kind == SHEAR) {
                if (m_zero) {
                    if (marks) TPCF_LAUNCH(SHEAR, true, true, true); else TPCF_LAUNCH(SHEAR, false, true, true);
                } else {
                    if (marks) TPCF_LAUNCH(SHEAR, true, false, true); else TPCF_LAUNCH(SHEAR, false, false, true);
                }
            } else {
                if (marks) TPCF_LAUNCH(KAPPA, true, true, true); else TPCF_LAUNCH(KAPPA, false, true, true);
            }
        } else {
            return fail(ctx, TPCF_ERR_UNSUPPORTED, "spherical geometry runs in FP64");
        }
    } else if (kind == SHEAR) {
        if (m_zero) {
            if (marks) TPCF_LAUNCH(SHEAR, true, true, false); else TPCF_LAUNCH(SHEAR, false, true, false);
        } else {
            if (marks) TPCF_LAUNCH(SHEAR, true, false, false); else TPCF_LAUNCH(SHEAR, false, false, false);
        }
    } else {
        if (marks) TPCF_LAUNCH(KAPPA, true, true, false); else TPCF_LAUNCH(KAPPA, false, true, false);
    }
#undef TPCF_LAUNCH
    TPCF_HIP(ctx, hipGetLastError());
    return TPCF_OK;
}

template int make_pair_args<float>(tpcf_ctx *, const tpcf_bins *, int, PairArgs<float> &, const float **);
template int make_pair_args<double>(tpcf_ctx *, const tpcf_bins *, int, PairArgs<double> &, const double **);
template CatArgs<float> cat_args<float>(const Prepared<float> &, int32_t, int32_t, double *, unsigned long long *,
                                        double *);
template CatArgs<double> cat_args<double>(const Prepared<double> &, int32_t, int32_t, double *, unsigned long long *,
                                          double *);
template int launch_pairs<float>(tpcf_ctx *, hipStream_t, int, bool, bool, const PairArgs<float> &,
                                 const CatArgs<float> &, const float *, const Item *, unsigned long long *);
template int launch_pairs<double>(tpcf_ctx *, hipStream_t, int, bool, bool, const PairArgs<double> &,
                                  const CatArgs<double> &, const double *, const Item *, unsigned long long *);
template int run_pairs<float>(tpcf_ctx *, const CallArgs<float> &, const Prepared<float> &, double *,
                              unsigned long long *, double *);
template int run_pairs<double>(tpcf_ctx *, const CallArgs<double> &, const Prepared<double> &, double *,
                               unsigned long long *, double *);

}  // namespace tpcf
