// batch.hip -- many catalogues as one graph: the field loop of the reference
// (src/main.f90:596-617, :639-660: one shear_2pcf / kappa_2pcf call per
// field, each writing its marks, :704-707) for mock catalogues of uneven
// sizes.
//
// tpcf_batch_create_* prepares every catalogue once (sorting into cells,
// segment, chunk and item tables: the positions and masks), optionally after
// an autotuning pass that times candidate cell sizes and work-item sizes on a
// sample of the catalogues.  tpcf_batch_run then, per catalogue, re-reads the
// values (val, w, z, e1/e2 or kappa, m) from the caller's device arrays into
// the prepared structure, zeroes the accumulators, runs the pair kernel,
// writes the per-lens marks into the caller's device arrays (when asked for),
// reduces and copies the per-bin sums to pinned host memory.  The first run
// captures all of that into one HIP graph whose catalogues are spread over
// kBranches parallel branches (so small catalogues run side by side); later
// runs replay the graph.  Device inputs only: the graph reads the same device
// pointers on every replay, so new values (a new mock realisation on the
// same positions) are picked up by the next run.
//
// With TPCF_SPHERICAL every catalogue is planned on the sphere as a single
// spherical call is (sphere.hip): FP64 projected positions and frames fixed
// at planning, FP64 copies of FP32 values redone at the start of each run.
#include <algorithm>
#include <cmath>
#include <memory>
#include <type_traits>
#include <vector>

#include "internal.hpp"

namespace tpcf {

namespace {

constexpr int kBranches = 8;

// One catalogue.  U is the precision the catalogue is computed in: the input
// precision on the flat sky, FP64 on the sphere.
template <typename U>
struct BatchCat {
    tpcf_catalog desc;
    tpcf_ctx *sub = nullptr;   // owns this catalogue's device buffers
    CallArgs<U> args;
    Prepared<U> p;
    PairArgs<U> A;
    double *pp = nullptr;
    unsigned long long *ppn = nullptr;
    double *ppo = nullptr;        // reduced [5][P_B][P_A][n_bin]
    long long *ppno = nullptr;    // reduced [P_B][P_A][n_bin]
    double *dtot = nullptr;       // [n_bin][5] FP64 sums, then [n_bin] int64 counts
    double *htot = nullptr;       // pinned copy of dtot
    double *hpp = nullptr;        // pinned copies of ppo / ppno (when asked for)
    long long *hppn = nullptr;
    int32_t *herr = nullptr;      // pinned copy of the preparation's error flags (mask breaches)
    int64_t pp_total = 0;
    bool want_pp = false;
    // spherical catalogues: the chord bins (args.kbins points here) and the
    // FP32 -> FP64 value copies each run repeats
    bool sph = false;
    tpcf_bins kb{};
    std::vector<F64Copy> copies;
    // per-lens marks: FP64 planes [n_bin][4][n_lens] and the caller's device
    // outputs xi_E, xi_B, K, w marks and marks_weight of mrb bytes per element
    bool marks = false;
    double *mk = nullptr;
    void *mout[5] = {};
    int mrb = 4;
};

// A bin group's pair arguments (bins, edges, tests) with one catalogue's work
// items: plan_cat fills only the item count and layout of the catalogue's A.
template <typename U>
PairArgs<U> with_items(const PairArgs<U> &group, const PairArgs<U> &planned)
{
    PairArgs<U> A = group;
    A.n_items = planned.n_items;
    A.per_xcd = planned.per_xcd;
    return A;
}

int sub_ctx(tpcf_ctx *parent, tpcf_ctx **out)
{
    tpcf_ctx *c = new (std::nothrow) tpcf_ctx();
    if (!c) return TPCF_ERR_ALLOC;
    c->device = parent->device;
    c->stream = parent->stream;
    c->opt_cell = parent->opt_cell;
    c->opt_cell_target = parent->opt_cell_target;
    c->opt_split = parent->opt_split;
    *out = c;
    return TPCF_OK;
}

void free_sub(tpcf_ctx *c)
{
    if (!c) return;
    for (auto &b : c->buf)
        if (b.p) (void)hipFree(b.p);
    delete c;
}

}  // namespace
}  // namespace tpcf

struct tpcf_batch {
    virtual ~tpcf_batch() = default;
    virtual int run() = 0;
    tpcf_batch_stats stats{};
};

namespace tpcf {
namespace {

template <typename U>
struct Batch final : tpcf_batch {
    tpcf_ctx *ctx = nullptr;   // used by create and run only: the destructor does not touch it
    int device = 0;
    tpcf_bins bins{};          // the caller's bins (the outputs' bin centres)
    tpcf_bins kb{};            // the bins the pair kernel compares with (chord edges on the sphere)
    bool sph = false;
    std::vector<std::unique_ptr<BatchCat<U>>> cats;
    // per bin group (kGroupBins bins per pair-kernel pass): the pair arguments
    // and the squared edges one per lane (the batch's own copies)
    std::vector<PairArgs<U>> A0;
    std::vector<const U *> e2lane;
    hipStream_t br[kBranches] = {};
    hipEvent_t fork = nullptr, join[kBranches] = {};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool use_graph = true;
    void *e2own = nullptr;   // the batch's own copy of the squared edges of every group

    // The pair arguments and edge tables of every bin group, in the batch's own memory.
    int make_groups()
    {
        const int ng = bin_groups(kb.n_bin);
        A0.resize(ng);
        e2lane.resize(ng);
        TPCF_HIP(ctx, hipMalloc(&e2own, sizeof(U) * kWave * ng));
        for (int g = 0; g < ng; ++g) {
            const U *tmp;
            TPCF_TRY(make_pair_args<U>(ctx, &kb, g, A0[g], &tmp));
            U *dst = static_cast<U *>(e2own) + (size_t)g * kWave;
            TPCF_HIP(ctx, hipMemcpyAsync(dst, tmp, sizeof(U) * kWave, hipMemcpyDeviceToDevice, ctx->stream));
            e2lane[g] = dst;
        }
        TPCF_HIP(ctx, hipStreamSynchronize(ctx->stream));
        return TPCF_OK;
    }

    ~Batch() override
    {
        // the context may already be gone: wait for the device instead of its
        // stream (a replay in flight must end before its graph is destroyed)
        (void)hipSetDevice(device);
        (void)hipDeviceSynchronize();
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (auto &c : cats) {
            if (c->htot) (void)hipHostFree(c->htot);
            if (c->hpp) (void)hipHostFree(c->hpp);
            if (c->hppn) (void)hipHostFree(c->hppn);
            if (c->herr) (void)hipHostFree(c->herr);
            free_sub(c->sub);
        }
        for (int i = 0; i < kBranches; ++i) {
            if (br[i]) (void)hipStreamDestroy(br[i]);
            if (join[i]) (void)hipEventDestroy(join[i]);
        }
        if (fork) (void)hipEventDestroy(fork);
        if (e2own) (void)hipFree(e2own);
    }

    // The work of one catalogue on stream s (kernel launches and copies).
    int enqueue(BatchCat<U> &c, hipStream_t s)
    {
        tpcf_ctx *sc = c.sub;
        const int n_bin = bins.n_bin;
        if (c.args.n_a == 0 || c.args.n_b == 0) {
            // no pairs: the marks outputs of the lenses are zero
            if (c.marks && c.args.n_a > 0)
                TPCF_TRY(launch_marks_out(sc, s, 0, n_bin, c.args.n_a, nullptr, nullptr, c.mout, c.mrb, false));
            return TPCF_OK;
        }
        for (const F64Copy &f : c.copies) launch_to_f64(s, f.src, f.dst, f.n);
        TPCF_TRY(refresh_values<U>(sc, c.args, c.p, s));
        if constexpr (sizeof(U) == 8) {
            // the re-gathered records hold 2 e1 w where the sphere keeps p
            if (c.sph) sph_records_p(s, c.p, c.args.sph_pb);
        }
        fill_f64(s, c.pp, kNQ * kSlots * c.pp_total);
        fill_u64(s, c.ppn, kSlots * c.pp_total);
        if (c.marks) fill_f64(s, c.mk, 4 * c.p.n_lens * n_bin);
        for (size_t g = 0; g < A0.size(); ++g) {
            const size_t g0 = g * kGroupBins;
            const CatArgs<U> C = cat_args<U>(c.p, c.args.n_pa, c.args.n_pb, c.pp + g0 * kNQ, c.ppn + g0,
                                             c.marks ? c.mk + g0 * 4 * (size_t)c.p.n_lens : nullptr);
            TPCF_TRY(launch_pairs<U>(sc, s, c.args.kind, c.p.m_zero, c.marks, with_items(A0[g], c.A), C, e2lane[g],
                                     c.p.items, nullptr));
        }
        if (c.marks)
            TPCF_TRY(launch_marks_out(sc, s, c.p.n_lens, n_bin, c.args.n_a, c.p.lidx, c.mk, c.mout, c.mrb,
                                      c.args.kind == KAPPA || c.p.m_zero));
        TPCF_TRY(launch_reduce(sc, s, c.args.n_pa, c.args.n_pb, n_bin, c.pp, c.ppn, c.ppo, c.ppno, c.dtot));
        TPCF_HIP(sc, hipMemcpyAsync(c.htot, c.dtot, sizeof(double) * n_bin * (kNQ + 1), hipMemcpyDeviceToHost, s));
        TPCF_HIP(sc, hipMemcpyAsync(c.herr, c.p.err, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        if (c.want_pp) {
            TPCF_HIP(sc, hipMemcpyAsync(c.hpp, c.ppo, sizeof(double) * kNQ * c.pp_total, hipMemcpyDeviceToHost, s));
            TPCF_HIP(sc, hipMemcpyAsync(c.hppn, c.ppno, sizeof(long long) * c.pp_total, hipMemcpyDeviceToHost, s));
        }
        return TPCF_OK;
    }

    // All catalogues, spread over the branches, joined back into ctx->stream.
    int enqueue_all()
    {
        hipStream_t s = ctx->stream;
        TPCF_HIP(ctx, hipEventRecord(fork, s));
        for (int i = 0; i < kBranches; ++i) TPCF_HIP(ctx, hipStreamWaitEvent(br[i], fork, 0));
        // largest first, dealt round-robin: the branches get similar work
        std::vector<int> order(cats.size());
        for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
        std::sort(order.begin(), order.end(), [&](int a, int b) {
            return (double)cats[a]->args.n_a * cats[a]->args.n_b > (double)cats[b]->args.n_a * cats[b]->args.n_b;
        });
        for (size_t k = 0; k < order.size(); ++k) {
            const int rc = enqueue(*cats[order[k]], br[k % kBranches]);
            if (rc) return rc;
        }
        for (int i = 0; i < kBranches; ++i) {
            TPCF_HIP(ctx, hipEventRecord(join[i], br[i]));
            TPCF_HIP(ctx, hipStreamWaitEvent(s, join[i], 0));
        }
        return TPCF_OK;
    }

    void write_outputs()
    {
        const int n_bin = bins.n_bin;
        int64_t total = 0;
        for (auto &cp : cats) {
            BatchCat<U> &c = *cp;
            tpcf_out *out = c.desc.out;
            if (!out) continue;
            const int rb = out->real_bytes == 8 ? 8 : 4;
            for (int b = 0; b < n_bin; ++b) {
                const double *t = c.args.n_a && c.args.n_b ? &c.htot[(size_t)b * kNQ] : nullptr;
                const double th = t ? t[0] : 0, gt = t ? t[1] : 0, gx = t ? t[2] : 0, w = t ? t[3] : 0,
                             k = t ? t[4] : 0;
                int64_t n = 0;   // the int64 counts follow the sums
                if (t) std::memcpy(&n, &c.htot[(size_t)n_bin * kNQ + b], sizeof n);
                total += n;
                auto put = [&](void *dst, double v) {
                    if (!dst) return;
                    if (rb == 8) static_cast<double *>(dst)[b] = v;
                    else static_cast<float *>(dst)[b] = (float)v;
                };
                put(out->theta, (double)bins.bin_centers[b]);
                put(out->mean_theta, th / w);   // NaN for an empty bin, src/tpcf.f90:369
                put(out->g_t, k > 0.0 ? gt / k : 0.0);
                put(out->g_x, k > 0.0 ? gx / k : 0.0);
                put(out->weights, w);
                if (out->n_pairs) out->n_pairs[b] = n;
                if (out->sum_theta) out->sum_theta[b] = th;
                if (out->sum_gt) out->sum_gt[b] = gt;
                if (out->sum_gx) out->sum_gx[b] = gx;
                if (out->sum_w) out->sum_w[b] = w;
                if (out->sum_k) out->sum_k[b] = k;
            }
            if (c.want_pp) {
                double *dst[kNQ] = {out->pp_theta, out->pp_gt, out->pp_gx, out->pp_w, out->pp_k};
                for (int q = 0; q < kNQ; ++q)
                    if (dst[q]) std::memcpy(dst[q], c.hpp + (size_t)q * c.pp_total, sizeof(double) * c.pp_total);
                if (out->pp_n) std::memcpy(out->pp_n, c.hppn, sizeof(long long) * c.pp_total);
            }
        }
        stats.n_pairs = total;
    }

    int run() override
    {
        TPCF_HIP(ctx, hipSetDevice(ctx->device));
        const double t0 = now_ms();
        hipStream_t s = ctx->stream;
        if (use_graph) {
            if (!exec) {
                TPCF_HIP(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                const int rc = enqueue_all();
                hipGraph_t g = nullptr;
                const hipError_t e = hipStreamEndCapture(s, &g);
                if (rc) {
                    if (g) (void)hipGraphDestroy(g);
                    return rc;
                }
                if (e != hipSuccess) return fail(ctx, TPCF_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
                graph = g;
                TPCF_HIP(ctx, hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
                hipGraphNode_t *nodes = nullptr;
                size_t n_nodes = 0;
                TPCF_HIP(ctx, hipGraphGetNodes(graph, nodes, &n_nodes));
                stats.graph_nodes = (int64_t)n_nodes;
            }
            TPCF_HIP(ctx, hipGraphLaunch(exec, s));
        } else {
            TPCF_TRY(enqueue_all());
        }
        TPCF_HIP(ctx, hipStreamSynchronize(s));
        // the masks are fixed at planning: a kept object that the new values
        // mask would still contribute, where the reference skips it per call
        for (size_t i = 0; i < cats.size(); ++i) {
            const BatchCat<U> &c = *cats[i];
            if (c.herr && (*c.herr & (8 | 16)))
                return fail(ctx, TPCF_ERR_MASK,
                            "catalogue " + std::to_string(i) + ": new values mask " +
                                ((*c.herr & 16) ? "a lens (w == 0 or z < 0)" : "a source (z < 0)") +
                                " that the batch was planned with; plan a new batch");
        }
        write_outputs();
        stats.ms_run = now_ms() - t0;
        stats.runs += 1;
        return TPCF_OK;
    }
};

// Plans one catalogue whose arrays have precision T into c (precision U:
// T on the flat sky, FP64 with sph).  kb: the batch's chord bins (sph only).
template <typename T, typename U>
int plan_cat(tpcf_ctx *ctx, const tpcf_bins *bins, const tpcf_bins *kb, bool sph, const tpcf_catalog &d,
             BatchCat<U> &c, bool keep_host)
{
    c.desc = d;
    const bool shear = (d.kind == SHEAR);
    if (d.kind != SHEAR && d.kind != KAPPA) return fail(ctx, TPCF_ERR_ARG, "catalogue kind must be SHEAR or KAPPA");
    if (d.n_A < 0 || d.n_B < 0 || d.n_A >= ((int64_t)1 << 31) || d.n_B >= ((int64_t)1 << 31))
        return fail(ctx, TPCF_ERR_ARG, "catalogue sizes must be in [0, 2^31)");
    if (d.n_A > 0 && (!d.x_A || !d.y_A || !d.val_A || !d.w_A || (shear && !d.z_A)))
        return fail(ctx, TPCF_ERR_ARG, "a required lens array is NULL");
    if (d.n_B > 0 && (!d.x_B || !d.y_B || !d.s1_B || !d.s2_B || !d.w_B || (shear && !d.z_B)))
        return fail(ctx, TPCF_ERR_ARG, "a required source array is NULL");
    if (d.out && d.out->real_bytes != 4 && d.out->real_bytes != 8)
        return fail(ctx, TPCF_ERR_ARG, "out->real_bytes must be 4 or 8");
    TPCF_TRY(sub_ctx(ctx, &c.sub));
    CallArgs<T> a;
    a.kind = d.kind;
    a.n_a = d.n_A;
    a.n_b = d.n_B;
    a.xa = (const T *)d.x_A; a.ya = (const T *)d.y_A; a.za = shear ? (const T *)d.z_A : nullptr;
    a.vala = (const T *)d.val_A; a.wa = (const T *)d.w_A;
    a.xb = (const T *)d.x_B; a.yb = (const T *)d.y_B; a.zb = shear ? (const T *)d.z_B : nullptr;
    a.s1b = (const T *)d.s1_B; a.s2b = (const T *)d.s2_B; a.wb = (const T *)d.w_B;
    a.mb = shear ? (const T *)d.m_B : nullptr;   // NULL: m == 0 for every source
    a.patch_a = d.patch_A;
    a.patch_b = d.patch_B;
    a.n_pa = d.patch_A ? d.n_patch_A : 1;
    a.n_pb = d.patch_B ? d.n_patch_B : 1;
    if (a.n_pa < 1 || a.n_pb < 1) return fail(ctx, TPCF_ERR_ARG, "n_patch must be >= 1");
    a.bins = bins;
    a.marks = false;
    a.verbose = false;
    if (d.out) {
        void *m[5] = {d.out->xi_E_marks, d.out->xi_B_marks, d.out->K_marks, d.out->w_marks, d.out->marks_weight};
        for (int q = 0; q < 5; ++q) {
            c.mout[q] = m[q];
            c.marks = c.marks || m[q] != nullptr;
        }
        c.mrb = d.out->real_bytes;
    }
    a.marks = c.marks;
    tpcf_ctx *sc = c.sub;
    if (sph) {
        if constexpr (sizeof(U) == 8) {
            c.sph = true;
            c.kb = *kb;
            // positions projected once; values copied to FP64 now and at each run
            TPCF_TRY(sph_setup<T>(sc, a, &c.kb, c.args, &c.copies));
        } else {
            return fail(ctx, TPCF_ERR_ARG, "spherical catalogues are planned in FP64");
        }
    } else {
        if constexpr (std::is_same<T, U>::value) c.args = a;
        else return fail(ctx, TPCF_ERR_ARG, "flat catalogues are planned in their input precision");
    }
    std::memset(&c.p, 0, sizeof c.p);
    c.pp_total = (int64_t)a.n_pa * a.n_pb * bins->n_bin;
    c.want_pp = d.out && (d.out->pp_n || d.out->pp_theta || d.out->pp_gt || d.out->pp_gx || d.out->pp_w || d.out->pp_k);
    if (a.n_a > 0 && a.n_b > 0) {
        TPCF_TRY(prepare<U>(sc, c.args, c.p));
        if constexpr (sizeof(U) == 8) {
            if (c.sph) TPCF_TRY(sph_finish(sc, c.args, c.p));
        }
        // m values are re-read on every run: the m == 0 kernel only when no m is given
        if (shear && a.mb) c.p.m_zero = false;
        const int64_t nblk = (c.p.n_items + kWavesPerBlock - 1) / kWavesPerBlock;
        c.A.n_items = c.p.n_items;
        c.A.per_xcd = (nblk + kSlots - 1) / kSlots;
        TPCF_TRY(ensure(sc, tpcf_ctx::B_PP, sizeof(double) * kNQ * kSlots * c.pp_total, (void **)&c.pp));
        TPCF_TRY(ensure(sc, tpcf_ctx::B_PPN, sizeof(unsigned long long) * kSlots * c.pp_total, (void **)&c.ppn));
        if (c.marks)
            TPCF_TRY(ensure(sc, tpcf_ctx::B_MARKS, sizeof(double) * 4 * std::max<int64_t>(c.p.n_lens, 1) * bins->n_bin,
                            (void **)&c.mk));
        void *vb;
        TPCF_TRY(ensure(sc, tpcf_ctx::B_OUT,
                        sizeof(double) * (kNQ + 1) * c.pp_total + sizeof(double) * bins->n_bin * (kNQ + 1) + 64, &vb));
        c.ppo = static_cast<double *>(vb);
        c.ppno = reinterpret_cast<long long *>(c.ppo + (size_t)kNQ * c.pp_total);
        c.dtot = reinterpret_cast<double *>(c.ppno + c.pp_total);
        // the sort's scratch is not needed any more
        for (int bidx : {(int)tpcf_ctx::B_SKEY, (int)tpcf_ctx::B_SKEY2, (int)tpcf_ctx::B_SIDX, (int)tpcf_ctx::B_LKEY,
                         (int)tpcf_ctx::B_LKEY2, (int)tpcf_ctx::B_LIDX, (int)tpcf_ctx::B_TMP, (int)tpcf_ctx::B_WORK,
                         (int)tpcf_ctx::B_SPH_PART}) {
            DevBuf &b = sc->buf[bidx];
            if (b.p) (void)hipFree(b.p);
            b.p = nullptr;
            b.cap = 0;
        }
    }
    if (keep_host) {
        TPCF_HIP(ctx, hipHostMalloc((void **)&c.htot, sizeof(double) * bins->n_bin * (kNQ + 1), hipHostMallocDefault));
        std::memset(c.htot, 0, sizeof(double) * bins->n_bin * (kNQ + 1));
        TPCF_HIP(ctx, hipHostMalloc((void **)&c.herr, sizeof(int32_t), hipHostMallocDefault));
        *c.herr = 0;
        if (c.want_pp) {
            TPCF_HIP(ctx, hipHostMalloc((void **)&c.hpp, sizeof(double) * kNQ * c.pp_total, hipHostMallocDefault));
            TPCF_HIP(ctx, hipHostMalloc((void **)&c.hppn, sizeof(long long) * c.pp_total, hipHostMallocDefault));
            std::memset(c.hpp, 0, sizeof(double) * kNQ * c.pp_total);
            std::memset(c.hppn, 0, sizeof(long long) * c.pp_total);
        }
    }
    return TPCF_OK;
}

// Times one eager run of the pair work of the given catalogues (ms), without
// their marks.
template <typename T, typename U>
int time_plan(tpcf_ctx *ctx, const tpcf_bins *bins, const tpcf_bins *kb, bool sph,
              const std::vector<const tpcf_catalog *> &sample, double *ms)
{
    Batch<U> b;
    b.ctx = ctx;
    b.device = ctx->device;
    b.bins = *bins;
    b.kb = *kb;
    b.sph = sph;
    b.use_graph = false;
    TPCF_TRY(b.make_groups());
    for (const tpcf_catalog *d : sample) {
        auto c = std::make_unique<BatchCat<U>>();
        tpcf_catalog dd = *d;
        dd.out = nullptr;
        TPCF_TRY((plan_cat<T, U>(ctx, &b.bins, &b.kb, sph, dd, *c, true)));
        b.cats.push_back(std::move(c));
    }
    TPCF_HIP(ctx, hipEventCreate(&b.fork));
    for (int i = 0; i < kBranches; ++i) {
        TPCF_HIP(ctx, hipStreamCreate(&b.br[i]));
        TPCF_HIP(ctx, hipEventCreate(&b.join[i]));
    }
    TPCF_TRY(b.run());   // warm
    hipEvent_t e0, e1;
    TPCF_HIP(ctx, hipEventCreate(&e0));
    TPCF_HIP(ctx, hipEventCreate(&e1));
    TPCF_HIP(ctx, hipEventRecord(e0, ctx->stream));
    TPCF_TRY(b.enqueue_all());
    TPCF_HIP(ctx, hipEventRecord(e1, ctx->stream));
    TPCF_HIP(ctx, hipEventSynchronize(e1));
    float f = 0;
    TPCF_HIP(ctx, hipEventElapsedTime(&f, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = f;
    return TPCF_OK;
}

template <typename T, typename U>
int batch_build(tpcf_ctx *ctx, int32_t n_cat, const tpcf_catalog *cats, const tpcf_bins *bins, uint32_t flags,
                tpcf_batch **out)
{
    const bool sph = (flags & TPCF_SPHERICAL) != 0;
    tpcf_bins kb = *bins;
    if (sph) TPCF_TRY(sph_chord_bins(ctx, bins, kb));
    const double t0 = now_ms();
    auto b = std::make_unique<Batch<U>>();
    b->ctx = ctx;
    b->device = ctx->device;
    b->bins = *bins;
    b->kb = kb;
    b->sph = sph;
    b->use_graph = !(flags & TPCF_BATCH_NO_GRAPH);
    const double saved_target = ctx->opt_cell_target, saved_split = ctx->opt_split;
    if ((flags & TPCF_BATCH_AUTOTUNE) && n_cat > 0) {
        // candidates: mean sources per cell x work-item size, timed on a
        // sample that spans the batch's sizes: per kind, up to kSample
        // catalogues evenly spaced in the list sorted by n_A n_B (timing only
        // the largest ones chose cells twice as large as the best for the
        // whole batch of mock fields: 15.3 s per run against 12.9 s)
        constexpr int kSample = 12;
        std::vector<const tpcf_catalog *> sample;
        for (int kind = 0; kind <= 1; ++kind) {
            std::vector<int> idx;
            for (int i = 0; i < n_cat; ++i)
                if (cats[i].kind == kind) idx.push_back(i);
            std::sort(idx.begin(), idx.end(), [&](int x, int y) {
                return (double)cats[x].n_A * cats[x].n_B > (double)cats[y].n_A * cats[y].n_B;
            });
            const int m = (int)idx.size(), s = std::min(m, kSample);
            for (int j = 0; j < s; ++j) sample.push_back(&cats[idx[s > 1 ? (int64_t)j * (m - 1) / (s - 1) : 0]]);
        }
        const double cand[][2] = {{32, 0.2}, {48, 0.25}, {64, 0.25}, {64, 0.35}, {128, 1.0}};
        double best = INFINITY;
        int bi = 3;
        const int nc = (int)(sizeof cand / sizeof cand[0]);
        for (int k = 0; k < nc; ++k) {
            ctx->opt_cell_target = cand[k][0];
            ctx->opt_split = cand[k][1];
            double ms = 0;
            const int rc = time_plan<T, U>(ctx, bins, &kb, sph, sample, &ms);
            if (rc) {
                ctx->opt_cell_target = saved_target;
                ctx->opt_split = saved_split;
                return rc;
            }
            b->stats.tune_ms[k] = ms;
            if (ms < best) {
                best = ms;
                bi = k;
            }
        }
        ctx->opt_cell_target = cand[bi][0];
        ctx->opt_split = cand[bi][1];
        b->stats.autotuned = 1;
    }
    b->stats.cell_target = cell_target<U>(ctx);
    b->stats.split = ctx->opt_split;
    // pair arguments and a private copy of the squared edges per bin group
    // (later calls on ctx reuse its own edge table)
    int rc = b->make_groups();
    for (int i = 0; i < n_cat && !rc; ++i) {
        auto c = std::make_unique<BatchCat<U>>();
        rc = plan_cat<T, U>(ctx, &b->bins, &b->kb, sph, cats[i], *c, true);
        if (!rc) b->stats.n_items += c->p.n_items;
        b->cats.push_back(std::move(c));
    }
    ctx->opt_cell_target = saved_target;
    ctx->opt_split = saved_split;
    if (rc) return rc;
    TPCF_HIP(ctx, hipEventCreate(&b->fork));
    for (int i = 0; i < kBranches; ++i) {
        TPCF_HIP(ctx, hipStreamCreate(&b->br[i]));
        TPCF_HIP(ctx, hipEventCreate(&b->join[i]));
    }
    TPCF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    b->stats.n_catalogs = n_cat;
    b->stats.ms_plan = now_ms() - t0;
    *out = b.release();
    return TPCF_OK;
}

template <typename T>
int batch_create(tpcf_ctx *ctx, int32_t n_cat, const tpcf_catalog *cats, const tpcf_bins *bins, uint32_t flags,
                 tpcf_batch **out)
{
    if (!ctx || !out) return TPCF_ERR_ARG;
    *out = nullptr;
    ctx->err.clear();
    if (n_cat < 0 || (n_cat > 0 && !cats)) return fail(ctx, TPCF_ERR_ARG, "bad catalogue list");
    TPCF_TRY(check_bins(ctx, bins));
    if (!(flags & TPCF_DEVICE_INPUTS))
        return fail(ctx, TPCF_ERR_ARG, "a batch reads device arrays: TPCF_DEVICE_INPUTS is required");
    if (flags & ~(uint32_t)(TPCF_DEVICE_INPUTS | TPCF_VERBOSE | TPCF_NO_KD_TREE | TPCF_BATCH_AUTOTUNE |
                            TPCF_BATCH_NO_GRAPH | TPCF_SPHERICAL | TPCF_DEVICE_MARKS))
        return fail(ctx, TPCF_ERR_ARG, "unknown flag");
    TPCF_HIP(ctx, hipSetDevice(ctx->device));
    if (flags & TPCF_SPHERICAL) return batch_build<T, double>(ctx, n_cat, cats, bins, flags, out);
    return batch_build<T, T>(ctx, n_cat, cats, bins, flags, out);
}

}  // namespace
}  // namespace tpcf

using namespace tpcf;

extern "C" {

int tpcf_batch_create_f32(tpcf_ctx *ctx, int32_t n_cat, const tpcf_catalog *cats, const tpcf_bins *bins,
                          uint32_t flags, tpcf_batch **batch)
{
    return batch_create<float>(ctx, n_cat, cats, bins, flags, batch);
}

int tpcf_batch_create_f64(tpcf_ctx *ctx, int32_t n_cat, const tpcf_catalog *cats, const tpcf_bins *bins,
                          uint32_t flags, tpcf_batch **batch)
{
    return batch_create<double>(ctx, n_cat, cats, bins, flags, batch);
}

int tpcf_batch_run(tpcf_batch *batch)
{
    if (!batch) return TPCF_ERR_ARG;
    return batch->run();
}

int tpcf_batch_get_stats(const tpcf_batch *batch, tpcf_batch_stats *stats)
{
    if (!batch || !stats) return TPCF_ERR_ARG;
    *stats = batch->stats;
    return TPCF_OK;
}

int tpcf_batch_destroy(tpcf_batch *batch)
{
    delete batch;
    return TPCF_OK;
}

}  // extern "C"
