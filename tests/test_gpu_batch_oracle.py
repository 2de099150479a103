"""The batched form (tpcf_batch_*) against the FP64 oracle, marks included.

The reference runs one correlation per field and writes each field's marks
(src/main.f90:596-617, :704-707); a batch must give, per catalogue, what the
oracle gives for that catalogue:

* a small batch of flat catalogues (kappa-kappa and shear, masks z < 0 and
  w = 0, m != 0, patches, an empty lens side), FP32 and FP64, graph replays
  with new values: pair counts per bin and per patch pair bit-exact; sums
  within 1e-6 (FP32; FP64: 1e-12) of each bin's scale (Theta, W, K
  relative, G against the sum of |terms|); marks within the same bars, the
  signed ones against the lens-bin sum of |terms|;
* seeded random batches (2 to 6 catalogues of either kind, 1 to 3000
  objects, their own sides and patch counts, m and masks or not) the same
  way;
* a spherical batch (FP32 and FP64 inputs) the same way, counts exact apart
  from pairs whose squared chord lies within 1e-9 of an edge (none here):
  shear with m, kappa-kappa, and shear without an m array (the m == 0
  kernel);
* stacking (src/main.f90:678-695): for fields far apart, the weight-stacked
  per-field results of a batch equal one correlation of the union (K = W for
  kappa-kappa and for shear with m = 0, so sum_f W_f g_f / sum_f W_f =
  sum G / sum K of the union), which pins the stacking formula;
* the full mocks_256 batch (256 fields, 512 catalogues of 5e4-2e6 objects,
  20 log bins 1'-100', 16 patches each): for 8 catalogues spread over the
  sizes, the marks of 256 sampled lenses (the full catalogue's cells,
  segments, chunks and work items) against the oracle's marks of those
  lenses; a 1024-lens subsample of each of those catalogues, run as extra
  catalogues of the same batch, against the oracle (counts and per patch pair
  counts bit-exact, sums 1e-6); the marks summed over all lenses equal each
  catalogue's per-bin sums (1e-9).
"""
import importlib
from types import SimpleNamespace

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
SHEAR, KAPPA = 0, 1
DEV = torch.device("cuda:0")


def host(d):
    return {k: v.detach().cpu().numpy() for k, v in d.items()}


def bound_marks(orc, kind, lens, src, edges, spherical=False):
    """Per lens and bin, sum over its pairs of |val| w_A w_B (|s1| + |s2|):
    a bound of the |terms| of the signed marks (|e1 c + e2 s| <= |e1| + |e2|),
    from an oracle kappa call.  It also counts the pairs a shear call skips by
    z, which only loosens it."""
    s1, s2 = ("e1", "e2") if kind == SHEAR else ("kappa_E", "kappa_B")
    L = dict(x=lens["x"], y=lens["y"], val=np.abs(lens["val"]), w=lens["w"])
    S = dict(x=src["x"], y=src["y"], kappa_E=np.abs(src[s1]) + np.abs(src[s2]), kappa_B=np.zeros(len(src["x"])),
             w=src["w"])
    return orc.correlate(KAPPA, L, S, edges, marks=True, spherical=spherical)["xi_e"]


def check_result(res, o, tol, patches=True):
    np.testing.assert_array_equal(res["n_pairs"], o["n_pairs"])
    for gk, ok in (("sum_theta", "theta"), ("sum_w", "w"), ("sum_k", "k")):
        den = np.where(o[ok] != 0, np.abs(o[ok]), 1.0)
        assert np.max(np.abs(res[gk] - o[ok]) / den) <= tol, gk
    for gk, ok, ak in (("sum_gt", "gt", "gt_abs"), ("sum_gx", "gx", "gx_abs")):
        den = np.where(o[ak] != 0, o[ak], 1.0)
        assert np.max(np.abs(res[gk] - o[ok]) / den) <= tol, gk
    if patches and "pp_n" in res:
        np.testing.assert_array_equal(res["pp_n"], o["pp_n"])
        for gk in ("pp_theta", "pp_w", "pp_k"):
            den = np.where(o[gk] != 0, np.abs(o[gk]), 1.0)
            assert np.max(np.abs(res[gk] - o[gk]) / den) <= tol, gk


def check_marks(res, o, bound, tol, rows=None):
    """GPU marks (n_bin, n_A) tensors against the oracle's [n][n_bin]."""
    n_bin, n_a = res["w_marks"].shape
    assert n_bin == o["w_marks"].shape[1], "marks must be (n_bin, n_A) views"
    for k in ("xi_E_marks", "xi_B_marks", "K_marks", "w_marks"):
        assert tuple(res[k].shape) == (n_bin, n_a), k
    assert tuple(res["marks_weight"].shape) == (n_a,)
    if rows is not None:   # checked here, so the device gather below stays in bounds
        rows = np.asarray(rows)
        assert rows.min() >= 0 and rows.max() < n_a
    sel = slice(None) if rows is None else torch.as_tensor(rows, device=res["w_marks"].device)
    g = {k: res[k][:, sel].T.cpu().numpy() for k in ("xi_E_marks", "xi_B_marks", "K_marks", "w_marks")}
    for gk, ok in (("K_marks", "k_marks"), ("w_marks", "w_marks")):
        den = np.where(o[ok] != 0, np.abs(o[ok]), 1.0)
        assert np.max(np.abs(g[gk] - o[ok]) / den) <= tol, gk
    for gk, ok in (("xi_E_marks", "xi_e"), ("xi_B_marks", "xi_b")):
        assert np.all(np.abs(g[gk] - o[ok]) <= tol * bound + 1e-300), gk
    mw = res["marks_weight"][sel].cpu().numpy()
    np.testing.assert_allclose(mw, o["marks_weight"], rtol=tol, atol=1e-300)


def flat_catalogue(seed, kind, n_a, n_b, side, n_patch, dtype, m=True, masks=True):
    r = np.random.default_rng(seed)
    lens = dict(x=r.uniform(0, side, n_a), y=r.uniform(0, side, n_a), val=r.normal(0, 0.02, n_a),
                w=r.uniform(0.5, 1.5, n_a), z=r.uniform(-0.1, 1.0, n_a) if masks else np.zeros(n_a))
    if masks:
        lens["w"][::23] = 0.0
    if kind == SHEAR:
        src = dict(x=r.uniform(0, side, n_b), y=r.uniform(0, side, n_b),
                   z=r.uniform(-0.1, 1.2, n_b) if masks else np.ones(n_b), e1=r.normal(0, 0.27, n_b),
                   e2=r.normal(0, 0.27, n_b), w=r.uniform(0.5, 1.5, n_b),
                   m=r.normal(0, 0.05, n_b) if m else np.zeros(n_b))
    else:
        lens.pop("z")
        src = dict(x=r.uniform(0, side, n_b), y=r.uniform(0, side, n_b), kappa_E=r.normal(0, 0.02, n_b),
                   kappa_B=r.normal(0, 0.02, n_b), w=r.uniform(0.5, 1.5, n_b))
    lens = {k: v.astype(dtype) for k, v in lens.items()}
    src = {k: v.astype(dtype) for k, v in src.items()}
    pa = r.integers(0, n_patch, n_a).astype(np.int32)
    pb = r.integers(0, n_patch, n_b).astype(np.int32)
    return dict(kind=kind, lens={k: torch.from_numpy(v).to(DEV) for k, v in lens.items()},
                src={k: torch.from_numpy(v).to(DEV) for k, v in src.items()},
                patch_A=torch.from_numpy(pa).to(DEV), patch_B=torch.from_numpy(pb).to(DEV), n_patch=n_patch)


def desc(c):
    synth = importlib.import_module("2pcf_amd.synth")
    return synth.batch_desc(c)


def oracle_of(orc, c, edges, marks=False, spherical=False):
    lens, src = host(c["lens"]), host(c["src"])
    if c["kind"] == SHEAR and "m" not in src:
        src["m"] = np.zeros(len(src["x"]))
    pa, pb = c["patch_A"].cpu().numpy(), c["patch_B"].cpu().numpy()
    return orc.correlate(c["kind"], lens, src, edges, pa, pb, c["n_patch"], c["n_patch"], marks=marks,
                         patches=True, spherical=spherical), lens, src


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_batch_marks_against_oracle(tp, orc, ctx, precision):
    dt = np.float32 if precision == "f32" else np.float64
    tol = 1e-6 if precision == "f32" else 1e-12
    bins = tp.create_bins(0.5, 25.0, 13, tp.LOG)
    cats = [flat_catalogue(800 + i, SHEAR if i % 2 else KAPPA, 2000 + 700 * i, 2500 + 500 * i, 80.0, 4, dt,
                           m=(i != 3)) for i in range(6)]
    cats.append(flat_catalogue(820, SHEAR, 1500, 1800, 80.0, 4, dt, masks=False, m=False))
    empty = flat_catalogue(821, SHEAR, 700, 10, 80.0, 4, dt)
    empty["src"] = {k: v[:0] for k, v in empty["src"].items()}
    empty["patch_B"] = empty["patch_B"][:0]
    cats.append(empty)
    b = tp.Batch([desc(c) for c in cats], bins, precision=precision, patches=True, marks=True, ctx=ctx)
    r = np.random.default_rng(4)
    for rep in range(2):
        print(f"\nbatch run {rep} ({precision})", flush=True)
        out = b.run()
        for i, (c, res) in enumerate(zip(cats, out)):
            if c["src"]["x"].numel() == 0:
                assert res["n_pairs"].sum() == 0
                assert tuple(res["w_marks"].shape) == (bins.n_bin, c["lens"]["x"].numel())
                assert float(res["w_marks"].abs().sum()) == 0.0
                continue
            o, lens, src = oracle_of(orc, c, bins.bin_edges, marks=True)
            assert o["n_pairs"].sum() > 0
            check_result(res, o, tol)
            check_marks(res, o, bound_marks(orc, c["kind"], lens, src, bins.bin_edges), tol)
        for c in cats:   # new values in place, masks unchanged
            c["lens"]["val"].copy_(torch.from_numpy(r.normal(0, 0.02, c["lens"]["val"].numel()).astype(dt)))
            for k in ("e1", "e2", "kappa_E"):
                if k in c["src"]:
                    c["src"][k].copy_(torch.from_numpy(r.normal(0, 0.2, c["src"][k].numel()).astype(dt)))
        torch.cuda.synchronize()
    b.close()


@pytest.mark.parametrize("seed", range(6))
def test_batch_random_configurations(tp, orc, ctx, seed):
    """Seeded random batches against the oracle: 2 to 6 catalogues of either
    kind, of 1 to 2500 lenses and 1 to 3000 sources, with their own field
    sides and patch counts, with or without m and masks, FP32 or FP64, over
    one set of bins, run twice."""
    r = np.random.default_rng(4000 + seed)
    dt = np.float32 if r.random() < 0.6 else np.float64
    precision = "f32" if dt == np.float32 else "f64"
    tol = 1e-6 if precision == "f32" else 1e-12
    bins = tp.create_bins(float(r.uniform(0.2, 1.0)), float(r.uniform(8.0, 30.0)), int(r.choice([3, 12, 20])),
                          tp.LOG)
    cats = []
    for i in range(int(r.integers(2, 7))):
        cats.append(flat_catalogue(4100 + 10 * seed + i, int(r.integers(0, 2)), int(r.choice([1, 64, 65, 500, 2500])),
                                   int(r.choice([1, 63, 700, 3000])), float(r.uniform(20.0, 90.0)),
                                   int(r.integers(1, 5)), dt, m=bool(r.random() < 0.5), masks=bool(r.random() < 0.5)))
    print(f"\nseed {seed}: {precision}, {len(cats)} catalogues, bins {bins.n_bin}", flush=True)
    b = tp.Batch([desc(c) for c in cats], bins, precision=precision, patches=True, marks=True, ctx=ctx)
    for rep in range(2):
        out = b.run()
        for c, res in zip(cats, out):
            o, lens, src = oracle_of(orc, c, bins.bin_edges, marks=True)
            check_result(res, o, tol)
            check_marks(res, o, bound_marks(orc, c["kind"], lens, src, bins.bin_edges), tol)
    b.close()


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_batch_spherical_against_oracle(tp, orc, ctx, precision):
    dt = np.float32 if precision == "f32" else np.float64
    deg = np.pi / 180.0
    bins = tp.create_bins(0.2 * deg, 6.0 * deg, 10, tp.LOG)
    cats = []
    for i, (lon0, lat0) in enumerate(((np.pi, 0.0), (0.0, 1.3), (0.05, -0.4))):
        c = flat_catalogue(840 + i, SHEAR if i != 1 else KAPPA, 1500, 2000, 1.0, 3, np.float64, m=(i == 0))
        r = np.random.default_rng(850 + i)
        for side in ("lens", "src"):
            n = c[side]["x"].numel()
            lon = np.mod(lon0 + r.uniform(-8, 8, n) * deg, 2 * np.pi)
            lat = np.clip(lat0 + r.uniform(-8, 8, n) * deg, -np.pi / 2, np.pi / 2)
            c[side]["x"] = torch.from_numpy(lon.astype(dt)).to(DEV)
            c[side]["y"] = torch.from_numpy(lat.astype(dt)).to(DEV)
        if i == 2:
            c["src"].pop("m")   # no m array: the batch takes the m == 0 kernel for this catalogue
        c["lens"] = {k: v.to(torch.float32 if dt == np.float32 else torch.float64) for k, v in c["lens"].items()}
        c["src"] = {k: v.to(torch.float32 if dt == np.float32 else torch.float64) for k, v in c["src"].items()}
        cats.append(c)
    torch.cuda.synchronize()
    b = tp.Batch([desc(c) for c in cats], bins, precision=precision, patches=True, marks=True, spherical=True,
                 ctx=ctx)
    for rep in range(2):
        out = b.run()
        for c, res in zip(cats, out):
            o, lens, src = oracle_of(orc, c, bins.bin_edges, marks=True, spherical=True)
            print(f"\nsphere batch run {rep}: pairs {int(res['n_pairs'].sum())}", flush=True)
            assert o["n_pairs"].sum() > 1000
            check_result(res, o, 1e-10)
            check_marks(res, o, bound_marks(orc, c["kind"], lens, src, bins.bin_edges, spherical=True), 1e-9)
        for c in cats:
            c["lens"]["val"].mul_(-1.5)
        torch.cuda.synchronize()
    b.close()


def test_batch_stacking_equals_union(tp, orc, ctx):
    """Fields 10^4 arcmin apart share no pair: the stacked per-field results
    (src/main.f90:678-695) equal the union's correlation when K = W."""
    tp_mod = importlib.import_module("2pcf_amd.tpcf")
    bins = tp.create_bins(1.0, 20.0, 8, tp.LOG)
    for kind in (KAPPA, SHEAR):
        cats = [flat_catalogue(870 + i, kind, 1500 + 300 * i, 1700, 60.0, 1, np.float32, m=False, masks=False)
                for i in range(4)]
        for i, c in enumerate(cats):
            for side in ("lens", "src"):
                c[side]["x"] += 1.0e4 * i
        torch.cuda.synchronize()
        b = tp.Batch([desc(c) for c in cats], bins, ctx=ctx)
        out = b.run()
        b.close()
        stacked = tp_mod.stack_fields([SimpleNamespace(**r) for r in out])
        lens = {k: np.concatenate([host(c["lens"])[k] for c in cats]) for k in cats[0]["lens"]}
        src = {k: np.concatenate([host(c["src"])[k] for c in cats]) for k in cats[0]["src"]}
        if kind == SHEAR:
            src["m"] = np.zeros(len(src["x"]))
        raw = orc.correlate(kind, lens, src, bins.bin_edges)
        o = orc.finalize(raw)
        np.testing.assert_array_equal(stacked["n_pairs"], raw["n_pairs"])
        for k in ("mean_theta", "weights"):
            np.testing.assert_allclose(stacked[k], o[k], rtol=2e-6, err_msg=k)
        # signed: within 2e-6 of the sum of |terms| / K (the FP32 path's bar, twice)
        for k, ak in (("g_t", "gt_abs"), ("g_x", "gx_abs")):
            assert np.all(np.abs(stacked[k] - o[k]) <= 2e-6 * raw[ak] / raw["k"]), k


def test_mocks_256_batch_against_oracle(tp, orc):
    synth = importlib.import_module("2pcf_amd.synth")
    cats = synth.mock_fields(256)
    torch.cuda.synchronize()
    n = len(cats)
    # 8 catalogues spread over the sizes (both kinds), by n_A n_B
    order = sorted(range(n), key=lambda i: cats[i]["lens"]["x"].numel() * cats[i]["src"]["x"].numel())
    chosen = [order[int(round(j * (n - 1) / 7))] for j in range(8)]
    if all(cats[i]["kind"] == cats[chosen[0]]["kind"] for i in chosen):
        chosen[-2] = order[-2]
    r = np.random.default_rng(256)
    extra = []
    for i in chosen:
        c = cats[i]
        na = c["lens"]["x"].numel()
        sel = torch.from_numpy(np.sort(r.choice(na, 1024, replace=False))).to(DEV)
        extra.append(dict(kind=c["kind"], lens={k: v[sel].contiguous() for k, v in c["lens"].items()},
                          src=c["src"], patch_A=c["patch_A"][sel].contiguous(), patch_B=c["patch_B"], n_patch=16))
    torch.cuda.synchronize()
    bins = tp.create_bins(1.0, 100.0, 20, tp.LOG)
    ctx = tp.Context(0)
    ctx.set_option(3, 64.0)    # the cells and items the mocks_256 benchmark's autotuner picks
    ctx.set_option(4, 0.25)
    b = tp.Batch([desc(c) for c in cats + extra], bins, patches=True, marks=set(chosen), ctx=ctx)
    out = b.run()
    st = b.stats
    print(f"\nmocks_256 + 8 samples: {st['n_catalogs']} catalogues, {st['n_pairs']} pairs, "
          f"run {st['ms_run']:.0f} ms, plan {st['ms_plan']:.0f} ms", flush=True)
    for j, i in enumerate(chosen):
        c, res = cats[i], out[i]
        lens, src = host(c["lens"]), host(c["src"])
        kind = c["kind"]
        if kind == SHEAR:
            src["m"] = np.zeros(len(src["x"]))
        print(f"catalogue {i}: kind {kind}, {len(lens['x'])} x {len(src['x'])}, pairs {int(res['n_pairs'].sum())}",
              flush=True)
        # marks of sampled lenses of the full catalogue
        rows = np.sort(r.choice(len(lens["x"]), 256, replace=False))
        ls = {k: v[rows] for k, v in lens.items()}
        o = orc.correlate(kind, ls, src, bins.bin_edges, marks=True)
        check_marks(res, o, bound_marks(orc, kind, ls, src, bins.bin_edges), 1e-6, rows=rows)
        # marks summed over the lenses = the per-bin sums
        for mk, sk in (("w_marks", "sum_w"), ("K_marks", "sum_k")):
            np.testing.assert_allclose(res[mk].sum(dim=1).cpu().numpy(), res[sk], rtol=1e-9)
        for mk, sk in (("xi_E_marks", "sum_gt"), ("xi_B_marks", "sum_gx")):
            tot = res[mk].sum(dim=1).cpu().numpy()
            assert np.all(np.abs(tot - res[sk]) <= 1e-9 * res[mk].abs().sum(dim=1).cpu().numpy()), mk
        # the 1024-lens subsample, a catalogue of the same batch
        e = extra[j]
        oe = orc.correlate(kind, host(e["lens"]), src, bins.bin_edges, e["patch_A"].cpu().numpy(),
                           c["patch_B"].cpu().numpy(), 16, 16, patches=True)
        assert oe["n_pairs"].sum() > 0
        check_result(out[n + j], oe, 1e-6)
    b.close()
    ctx.close()
