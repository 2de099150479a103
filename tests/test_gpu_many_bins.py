"""More bins than one pair-kernel pass takes (63: one squared edge per
wavefront lane).  The reference takes any --n-bin (src/main.f90:354-357); a
call with more bins runs one pass per group of up to 63 consecutive bins, a
pair at an inner group's top edge belonging to the next group, exactly as in
a single pass.  Against the FP64 oracle:

* shear FP32, 150 log bins (three groups), masks, m != 0, patches and marks:
  counts per bin and per patch pair bit-exact, sums within 1e-6 (the FP32
  bar), marks within 1e-6 of each lens-bin's sum of |terms|;
* kappa-kappa auto-correlation FP64, 100 linear bins from x_min = 0 (the
  zero-separation self pairs in bin 0 of the first group), integer-lattice
  positions so that many pairs sit exactly on edges, group boundaries
  included: counts bit-exact, sums within 1e-12;
* spherical shear, 70 log bins: counts exact (no pair within 1e-9 of an
  edge's squared chord in this catalogue), sums within 1e-10;
* a batch of catalogues with 90 bins and marks, and the bootstrap with 100
  bins, against the oracle; the adjoint refuses more than 63 bins.
"""
import importlib

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
SHEAR, KAPPA = 0, 1
DEV = torch.device("cuda:0")


def check(g, o, tol, patches=False):
    np.testing.assert_array_equal(g.n_pairs, o["n_pairs"])
    for gk, ok in (("sum_theta", "theta"), ("sum_w", "w"), ("sum_k", "k")):
        den = np.where(o[ok] != 0, np.abs(o[ok]), 1.0)
        assert np.max(np.abs(getattr(g, gk) - o[ok]) / den) <= tol, gk
    for gk, ok, ak in (("sum_gt", "gt", "gt_abs"), ("sum_gx", "gx", "gx_abs")):
        den = np.where(o[ak] != 0, o[ak], 1.0)
        assert np.max(np.abs(getattr(g, gk) - o[ok]) / den) <= tol, gk
    if patches:
        np.testing.assert_array_equal(g.pp_n, o["pp_n"])
        for gk in ("pp_theta", "pp_w", "pp_k"):
            den = np.where(o[gk] != 0, np.abs(o[gk]), 1.0)
            assert np.max(np.abs(getattr(g, gk) - o[gk]) / den) <= tol, gk


def bound_marks(orc, lens, src, edges, s1, s2, spherical=False):
    L = dict(x=lens["x"], y=lens["y"], val=np.abs(lens["val"]), w=lens["w"])
    S = dict(x=src["x"], y=src["y"], kappa_E=np.abs(src[s1]) + np.abs(src[s2]), kappa_B=np.zeros(len(src["x"])),
             w=src["w"])
    return orc.correlate(KAPPA, L, S, edges, marks=True, spherical=spherical)["xi_e"]


def test_shear_f32_150_bins_marks_patches(tp, orc, ctx):
    r = np.random.default_rng(150)
    n_a, n_b, side = 4000, 5000, 120.0
    f = np.float32
    lens = dict(x=r.uniform(0, side, n_a).astype(f), y=r.uniform(0, side, n_a).astype(f),
                z=r.uniform(-0.1, 1.0, n_a).astype(f), val=r.normal(0, 0.02, n_a).astype(f),
                w=r.uniform(0.5, 1.5, n_a).astype(f))
    lens["w"][::31] = 0.0
    src = dict(x=r.uniform(0, side, n_b).astype(f), y=r.uniform(0, side, n_b).astype(f),
               z=r.uniform(-0.1, 1.2, n_b).astype(f), e1=r.normal(0, 0.27, n_b).astype(f),
               e2=r.normal(0, 0.27, n_b).astype(f), w=r.uniform(0.5, 1.5, n_b).astype(f),
               m=r.normal(0, 0.05, n_b).astype(f))
    pa = r.integers(0, 5, n_a).astype(np.int32)
    pb = r.integers(0, 5, n_b).astype(np.int32)
    bins = tp.create_bins(0.3, 60.0, 150, tp.LOG)
    g = tp.correlate(SHEAR, lens, src, bins, patch_A=pa, patch_B=pb, n_patch_A=5, n_patch_B=5, real_bytes=8,
                     marks=True, ctx=ctx)
    o = orc.correlate(SHEAR, lens, src, bins.bin_edges, pa, pb, 5, 5, marks=True, patches=True)
    print("\n150 bins: pairs", int(g.n_pairs.sum()), "empty bins", int(np.sum(o["n_pairs"] == 0)), flush=True)
    assert np.sum(o["n_pairs"] > 0) >= 140
    check(g, o, 1e-6, patches=True)
    bound = bound_marks(orc, lens, src, bins.bin_edges, "e1", "e2")
    for gk, ok in (("xi_E_marks", "xi_e"), ("xi_B_marks", "xi_b")):
        assert np.all(np.abs(np.asarray(getattr(g, gk)).T - o[ok]) <= 1e-6 * bound + 1e-300), gk
    for gk, ok in (("K_marks", "k_marks"), ("w_marks", "w_marks")):
        a = np.asarray(getattr(g, gk)).T
        den = np.where(o[ok] != 0, np.abs(o[ok]), 1.0)
        assert np.max(np.abs(a - o[ok]) / den) <= 1e-6, gk
    np.testing.assert_allclose(g.marks_weight, o["marks_weight"], rtol=1e-6)


def test_kappa_f64_100_lin_bins_from_zero_lattice(tp, orc, ctx):
    r = np.random.default_rng(100)
    n = 3000
    # integer lattice: many separations are exact integers, so pairs sit on
    # the linear edges 0, 0.3, ..., and on the group boundaries 18.9 and 37.8
    x = r.integers(0, 60, n).astype(np.float64)
    y = r.integers(0, 60, n).astype(np.float64)
    val = r.normal(0, 0.02, n)
    w = r.uniform(0.5, 1.5, n)
    lens = dict(x=x, y=y, val=val, w=w)
    src = dict(x=x, y=y, kappa_E=val, kappa_B=r.normal(0, 0.02, n), w=w)
    bins = tp.create_bins(0.0, 30.0, 100, tp.LIN)
    g = tp.correlate(KAPPA, lens, src, bins, precision="f64", real_bytes=8, ctx=ctx)
    o = orc.correlate(KAPPA, lens, src, bins.bin_edges)
    print("\n100 lin bins from 0: pairs", int(g.n_pairs.sum()), "bin 0:", int(g.n_pairs[0]), flush=True)
    assert o["n_pairs"][0] >= n          # the self pairs (d = 0) and coincident points
    check(g, o, 1e-12)


def test_sphere_70_bins(tp, orc, ctx):
    r = np.random.default_rng(70)
    deg = np.pi / 180
    n_a, n_b = 2500, 3000
    lon = lambda k: np.mod(np.pi + r.uniform(-10, 10, k) * deg, 2 * np.pi)   # noqa: E731
    lat = lambda k: r.uniform(-10, 10, k) * deg                              # noqa: E731
    lens = dict(x=lon(n_a), y=lat(n_a), z=np.zeros(n_a), val=r.normal(0, 0.02, n_a), w=r.uniform(0.5, 1.5, n_a))
    src = dict(x=lon(n_b), y=lat(n_b), z=np.ones(n_b), e1=r.normal(0, 0.27, n_b), e2=r.normal(0, 0.27, n_b),
               w=r.uniform(0.5, 1.5, n_b), m=np.zeros(n_b))
    bins = tp.create_bins(0.1 * deg, 8.0 * deg, 70, tp.LOG)
    g = tp.correlate(SHEAR, lens, src, bins, precision="f64", real_bytes=8, ctx=ctx, spherical=True)
    o = orc.correlate(SHEAR, lens, src, bins.bin_edges, spherical=True)
    # device and host sincos may differ in the last bit: a pair whose squared
    # chord lies within 1e-9 of an edge's may fall on either side
    unit = lambda lo, la: np.stack([np.cos(la) * np.cos(lo), np.cos(la) * np.sin(lo), np.sin(la)], 1)  # noqa: E731
    ua, ub = unit(lens["x"], lens["y"]), unit(src["x"], src["y"])
    e2c = orc.chord_edges_squared(bins.bin_edges)
    amb = 0
    for i0 in range(0, n_a, 256):
        d = ua[i0:i0 + 256, None, :] - ub[None, :, :]
        c2 = np.einsum("ijk,ijk->ij", d, d).ravel()
        amb += int(np.sum(np.min(np.abs(c2[:, None] - e2c[None, :]) / e2c[None, :], axis=1) <= 1e-9))
    print("\nsphere 70 bins: pairs", int(g.n_pairs.sum()), "near-edge pairs", amb, flush=True)
    assert np.sum(np.abs(g.n_pairs - o["n_pairs"])) <= amb
    if amb == 0:
        check(g, o, 1e-10)


def test_batch_90_bins_marks(tp, orc, ctx):
    r = np.random.default_rng(90)
    bins = tp.create_bins(0.5, 30.0, 90, tp.LOG)
    cats, descs = [], []
    for i, kind in enumerate((KAPPA, SHEAR)):
        n_a, n_b = 2500 + 500 * i, 3000
        f = np.float32
        lens = dict(x=r.uniform(0, 90, n_a).astype(f), y=r.uniform(0, 90, n_a).astype(f),
                    val=r.normal(0, 0.02, n_a).astype(f), w=r.uniform(0.5, 1.5, n_a).astype(f))
        if kind == SHEAR:
            lens["z"] = np.zeros(n_a, f)
            src = dict(x=r.uniform(0, 90, n_b).astype(f), y=r.uniform(0, 90, n_b).astype(f), z=np.ones(n_b, f),
                       e1=r.normal(0, 0.27, n_b).astype(f), e2=r.normal(0, 0.27, n_b).astype(f),
                       w=r.uniform(0.5, 1.5, n_b).astype(f), m=np.zeros(n_b, f))
        else:
            src = dict(x=r.uniform(0, 90, n_b).astype(f), y=r.uniform(0, 90, n_b).astype(f),
                       kappa_E=r.normal(0, 0.02, n_b).astype(f), kappa_B=r.normal(0, 0.02, n_b).astype(f),
                       w=r.uniform(0.5, 1.5, n_b).astype(f))
        dl = {k: torch.from_numpy(v).to(DEV) for k, v in lens.items()}
        ds = {k: torch.from_numpy(v).to(DEV) for k, v in src.items()}
        cats.append((kind, lens, src, dl, ds))
        descs.append(dict(kind=kind, lens={k: v.data_ptr() for k, v in dl.items()},
                          src={k: v.data_ptr() for k, v in ds.items()}, n_A=n_a, n_B=n_b))
    torch.cuda.synchronize()
    b = tp.Batch(descs, bins, marks=True, ctx=ctx)
    out = b.run()
    for (kind, lens, src, _, _), res in zip(cats, out):
        o = orc.correlate(kind, lens, src, bins.bin_edges, marks=True)
        np.testing.assert_array_equal(res["n_pairs"], o["n_pairs"])
        den = np.where(o["w"] != 0, np.abs(o["w"]), 1.0)
        assert np.max(np.abs(res["sum_w"] - o["w"]) / den) <= 1e-6
        assert tuple(res["w_marks"].shape) == (90, len(lens["x"]))
        wm = res["w_marks"].T.cpu().numpy()
        den = np.where(o["w_marks"] != 0, np.abs(o["w_marks"]), 1.0)
        assert np.max(np.abs(wm - o["w_marks"]) / den) <= 1e-6
        s1, s2 = ("e1", "e2") if kind == SHEAR else ("kappa_E", "kappa_B")
        bound = bound_marks(orc, lens, src, bins.bin_edges, s1, s2)
        xe = res["xi_E_marks"].T.cpu().numpy()
        assert np.all(np.abs(xe - o["xi_e"]) <= 1e-6 * bound + 1e-300)
    b.close()


def test_bootstrap_100_bins(tp, orc, ctx):
    r = np.random.default_rng(101)
    P, nb = 12, 100
    k = r.uniform(1, 5, (P, nb))
    blocks = dict(g_t=r.normal(0, 0.01, (P, nb)) * k, g_x=r.normal(0, 0.01, (P, nb)) * k, k=k,
                  weight=r.uniform(1, 2, P))
    M = r.integers(0, 3, (70, P)).astype(np.int32)
    g = tp.bootstrap(blocks, 70, multiplicity=M, ctx=ctx)
    o = orc.bootstrap(blocks, M)
    for key in ("g_t", "g_x"):
        np.testing.assert_allclose(g[key], o[key], rtol=1e-12, atol=1e-15)
    c = orc.covariance(o["g_t"])
    assert np.max(np.abs(g["cov_t"] - c)) <= 1e-12 * np.max(np.abs(c))


def test_adjoint_refuses_more_than_63_bins(tp, ctx):
    r = np.random.default_rng(64)
    f = np.float32
    lens = dict(x=r.uniform(0, 20, 50).astype(f), y=r.uniform(0, 20, 50).astype(f), z=np.zeros(50, f),
                val=np.ones(50, f), w=np.ones(50, f))
    src = dict(x=lens["x"], y=lens["y"], z=np.ones(50, f), e1=np.zeros(50, f), e2=np.zeros(50, f), w=np.ones(50, f),
               m=np.zeros(50, f))
    bins = tp.create_bins(1.0, 10.0, 64, tp.LOG)
    with pytest.raises(tp.TpcfError) as e:
        tp.adjoint(SHEAR, lens, src, bins, np.ones(64), np.ones(64), ctx=ctx)
    assert e.value.code == tp._native.ERR_UNSUPPORTED
