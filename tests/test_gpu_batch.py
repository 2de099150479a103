"""Many catalogues as one HIP graph (tpcf_batch_*) on the GPU.

Each catalogue of a batch must give what a single correlation call gives on
the same inputs (pair counts and per-patch-pair counts identical; FP64 sums
equal up to the rounding of FP64 atomics in another order, 1e-12 of the bin's
scale).  The graph is replayed several times with new values written into the
same device arrays between replays (and eager work in between); every replay
is checked against eager single calls on the values of that moment.  The
eager (no graph) mode and the autotuner are checked the same way."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
SHEAR, KAPPA = 0, 1
DEV = torch.device("cuda:0")


def make_cat(seed, kind, n_a, n_b, side, n_patch=1, m=True):
    r = np.random.default_rng(seed)
    f = np.float32
    lens = dict(x=r.uniform(0, side, n_a).astype(f), y=r.uniform(0, side, n_a).astype(f),
                val=r.normal(0, 0.02, n_a).astype(f), w=r.uniform(0.5, 1.5, n_a).astype(f))
    if kind == SHEAR:
        lens["z"] = r.uniform(-0.1, 1.0, n_a).astype(f)
        src = dict(x=r.uniform(0, side, n_b).astype(f), y=r.uniform(0, side, n_b).astype(f),
                   z=r.uniform(-0.1, 1.2, n_b).astype(f), e1=r.normal(0, 0.27, n_b).astype(f),
                   e2=r.normal(0, 0.27, n_b).astype(f), w=r.uniform(0.5, 1.5, n_b).astype(f))
        if m:
            src["m"] = r.normal(0, 0.02, n_b).astype(f)
    else:
        src = dict(x=r.uniform(0, side, n_b).astype(f), y=r.uniform(0, side, n_b).astype(f),
                   kappa_E=r.normal(0, 0.02, n_b).astype(f), kappa_B=r.normal(0, 0.02, n_b).astype(f),
                   w=r.uniform(0.5, 1.5, n_b).astype(f))
    c = dict(kind=kind, host_lens=lens, host_src=src)
    c["dl"] = {k: torch.from_numpy(v).to(DEV) for k, v in lens.items()}
    c["ds"] = {k: torch.from_numpy(v).to(DEV) for k, v in src.items()}
    if n_patch > 1:
        c["pa"] = torch.from_numpy(r.integers(0, n_patch, n_a).astype(np.int32)).to(DEV)
        c["pb"] = torch.from_numpy(r.integers(0, n_patch, n_b).astype(np.int32)).to(DEV)
        c["n_patch"] = n_patch
    return c


def desc(c):
    d = dict(kind=c["kind"], lens={k: v.data_ptr() for k, v in c["dl"].items()},
             src={k: v.data_ptr() for k, v in c["ds"].items()}, n_A=c["dl"]["x"].numel(), n_B=c["ds"]["x"].numel())
    if "pa" in c:
        d.update(patch_A=c["pa"].data_ptr(), patch_B=c["pb"].data_ptr(), n_patch_A=c["n_patch"],
                 n_patch_B=c["n_patch"])
    return d


def single(tp, ctx, c, bins):
    lens = {k: v.cpu().numpy() for k, v in c["dl"].items()}
    src = {k: v.cpu().numpy() for k, v in c["ds"].items()}
    if c["kind"] == SHEAR and "m" not in src:
        src["m"] = np.zeros(len(src["x"]), np.float32)
    kw = {}
    if "pa" in c:
        kw = dict(patch_A=c["pa"].cpu().numpy(), patch_B=c["pb"].cpu().numpy(), n_patch_A=c["n_patch"],
                  n_patch_B=c["n_patch"])
    return tp.correlate(c["kind"], lens, src, bins, real_bytes=8, ctx=ctx, **kw)


def compare(res, g, patches):
    np.testing.assert_array_equal(res["n_pairs"], g.n_pairs)
    for k, gk in (("sum_theta", "sum_theta"), ("sum_gt", "sum_gt"), ("sum_gx", "sum_gx"), ("sum_w", "sum_w"),
                  ("sum_k", "sum_k")):
        a, b = res[k], getattr(g, gk)
        assert np.all(np.abs(a - b) <= 1e-12 * np.abs(b).max() + 1e-300), k
    if patches:
        np.testing.assert_array_equal(res["pp_n"], g.pp_n)
        for k in ("pp_theta", "pp_gt", "pp_w"):
            a, b = res[k], getattr(g, k)
            assert np.all(np.abs(a - b) <= 1e-12 * np.abs(b).max() + 1e-300), k


def catalogue_set():
    r = np.random.default_rng(77)
    cats = []
    for i in range(10):
        kind = SHEAR if i % 2 else KAPPA
        n_a = int(np.exp(r.uniform(np.log(300), np.log(6000))))
        n_b = int(np.exp(r.uniform(np.log(300), np.log(6000))))
        cats.append(make_cat(500 + i, kind, n_a, n_b, 60.0, n_patch=4 if i % 3 == 0 else 1, m=(i % 4 != 1)))
    cats.append(make_cat(600, SHEAR, 0, 500, 60.0))      # empty lens side
    return cats


@pytest.mark.parametrize("graph", [True, False])
def test_batch_matches_single_calls_and_replays(tp, ctx, graph):
    bins = tp.create_bins(0.5, 20.0, 12, tp.LOG)
    cats = catalogue_set()
    b = tp.Batch([desc(c) for c in cats], bins, graph=graph, patches=True, ctx=ctx)
    r = np.random.default_rng(5)
    for rep in range(3):
        print(f"\nbatch run {rep} (graph {graph})", flush=True)
        out = b.run()
        for c, res in zip(cats, out):
            compare(res, single(tp, ctx, c, bins), "pa" in c)
        # new values in place (same positions, weights and masks), eager work in between
        for c in cats:
            for k, v in c["dl"].items():
                if k == "val":
                    v.copy_(torch.from_numpy(r.normal(0, 0.02, v.numel()).astype(np.float32)).to(DEV))
            for k, v in c["ds"].items():
                if k in ("e1", "e2", "kappa_E", "kappa_B"):
                    v.copy_(torch.from_numpy(r.normal(0, 0.2, v.numel()).astype(np.float32)).to(DEV))
        torch.cuda.synchronize()
    st = b.stats
    print("stats:", st, flush=True)
    assert st["runs"] == 3
    if graph:
        assert st["graph_nodes"] > 0
    b.close()


def test_batch_cell_target_reported(tp, ctx):
    """Without the autotuner a batch plans its catalogues with the automatic
    cell size of single calls: 128 mean sources per cell for FP32 geometry,
    64 for FP64, or TPCF_OPT_CELL_TARGET of its context; stats report it."""
    bins = tp.create_bins(1.0, 30.0, 10, tp.LOG)
    cats = [make_cat(720 + i, KAPPA if i % 2 else SHEAR, 3000, 4000, 80.0) for i in range(2)]
    # an FP64 batch reads FP64 arrays
    cats64 = [dict(c, dl={k: v.double() for k, v in c["dl"].items()}, ds={k: v.double() for k, v in c["ds"].items()})
              for c in cats]
    for precision, cs, want in (("f32", cats, 128.0), ("f64", cats64, 64.0)):
        b = tp.Batch([desc(c) for c in cs], bins, precision=precision, ctx=ctx)
        assert b.stats["autotuned"] == 0 and b.stats["cell_target"] == want, b.stats
        b.close()
    c1 = tp.Context(0)
    c1.set_option(3, 40.0)
    b = tp.Batch([desc(c) for c in cats], bins, ctx=c1)
    assert b.stats["cell_target"] == 40.0, b.stats
    b.close()
    c1.close()


def test_batch_autotune(tp, ctx):
    bins = tp.create_bins(1.0, 30.0, 10, tp.LOG)
    cats = [make_cat(700 + i, KAPPA if i % 2 else SHEAR, 4000, 5000, 80.0) for i in range(4)]
    b = tp.Batch([desc(c) for c in cats], bins, autotune=True, ctx=ctx)
    st = b.stats
    print("autotune:", st, flush=True)
    assert st["autotuned"] == 1 and all(t > 0 for t in st["tune_ms"])
    # the candidates of tpcf_batch_create_* (batch.hip), in order
    cands = [(32, .2), (48, .25), (64, .25), (64, .35), (128, 1.)]
    assert st["tune_ms"][cands.index((st["cell_target"], st["split"]))] == min(st["tune_ms"])
    out = b.run()
    # single calls with the settings the autotuner chose (the same cells and
    # segments, so the same FP32 partial sums)
    c1 = tp.Context(0)
    c1.set_option(3, st["cell_target"])
    c1.set_option(4, st["split"])
    for c, res in zip(cats, out):
        compare(res, single(tp, c1, c, bins), False)
    c1.close()
    b.close()


def test_batch_outlives_context(tp):
    """A batch destroyed after its context (a test failure or a garbage
    collector can order them so) frees its own resources without touching
    the context."""
    bins = tp.create_bins(1.0, 30.0, 10, tp.LOG)
    c0 = tp.Context(0)
    b = tp.Batch([desc(make_cat(720, KAPPA, 2000, 2500, 60.0))], bins, ctx=c0)
    b.run()
    c0.close()
    b.close()
