"""Parity of the HIP path (through the C-ABI) with the CPU FP64 oracle.

Bar: pair counts per bin and per patch pair bit-exact; positive sums (Theta, W,
K) within 1e-6 relative in every bin (FP32 path; 1e-12 for FP64); signed sums
(G_t, G_x, xi_E, xi_B) within 1e-6 of the bin's sum of |terms| (a signed sum of
zero-mean terms can cancel, so its plain relative error is ill-conditioned).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHEAR, KAPPA = 0, 1


def _rng(seed):
    return np.random.default_rng(seed)


def uniform_case(seed, n_a, n_b, side, kind, z=False, zero_w=False):
    r = _rng(seed)
    f = np.float32
    lens = dict(x=r.uniform(0, side, n_a).astype(f), y=r.uniform(0, side, n_a).astype(f),
                val=r.normal(0, 0.02, n_a).astype(f), w=r.uniform(0.5, 1.5, n_a).astype(f),
                z=(r.uniform(-0.2, 1.0, n_a) if z else np.zeros(n_a)).astype(f))
    if zero_w:
        lens["w"][r.random(n_a) < 0.1] = 0
    if kind == SHEAR:
        src = dict(x=r.uniform(0, side, n_b).astype(f), y=r.uniform(0, side, n_b).astype(f),
                   z=(r.uniform(-0.2, 1.2, n_b) if z else np.ones(n_b)).astype(f),
                   e1=r.normal(0, 0.27, n_b).astype(f), e2=r.normal(0, 0.27, n_b).astype(f),
                   w=r.uniform(0.5, 1.5, n_b).astype(f), m=r.normal(0, 0.02, n_b).astype(f))
    else:
        src = dict(x=r.uniform(0, side, n_b).astype(f), y=r.uniform(0, side, n_b).astype(f),
                   kappa_E=r.normal(0, 0.02, n_b).astype(f), kappa_B=r.normal(0, 0.02, n_b).astype(f),
                   w=r.uniform(0.5, 1.5, n_b).astype(f))
    if zero_w:
        src["w"][r.random(n_b) < 0.1] = 0
    return lens, src


def check(tp, orc, kind, lens, src, bins, *, precision="f32", marks=False, patch_A=None, patch_B=None,
          n_pa=1, n_pb=1, ctx=None, tol=None):
    if tol is None:
        tol = 1e-6 if precision == "f32" else 1e-12
    g = tp.correlate(kind, lens, src, bins, precision=precision, marks=marks, patch_A=patch_A, patch_B=patch_B,
                     n_patch_A=n_pa, n_patch_B=n_pb, real_bytes=8, ctx=ctx)
    patches = patch_A is not None or patch_B is not None
    o = orc.correlate(kind, lens, src, bins.bin_edges, patch_A, patch_B, n_pa, n_pb, marks=marks, patches=patches)
    print("\npairs per bin:", g.n_pairs.tolist())
    print("stats:", {k: g.stats[k] for k in ("n_chunks", "n_segments", "cell_size", "ms_pairs")})
    np.testing.assert_array_equal(g.n_pairs, o["n_pairs"])
    errs = {}
    for gk, ok in (("sum_theta", "theta"), ("sum_w", "w"), ("sum_k", "k")):
        den = np.where(o[ok] != 0, np.abs(o[ok]), 1.0)
        errs[gk] = float(np.max(np.abs(g.__dict__[gk] - o[ok]) / den))
    for gk, ok, ak in (("sum_gt", "gt", "gt_abs"), ("sum_gx", "gx", "gx_abs")):
        den = np.where(o[ak] != 0, o[ak], 1.0)
        errs[gk] = float(np.max(np.abs(g.__dict__[gk] - o[ok]) / den))
    print("max errors:", errs)
    for k, v in errs.items():
        assert v <= tol, (k, v)
    # normalised outputs follow from the sums (src/tpcf.f90:369-376)
    with np.errstate(invalid="ignore", divide="ignore"):
        mt = o["theta"] / o["w"]
    np.testing.assert_array_equal(np.isnan(g.mean_theta), np.isnan(mt))
    if patches:
        np.testing.assert_array_equal(g.pp_n, o["pp_n"])
        for gk, ok in (("pp_theta", "pp_theta"), ("pp_w", "pp_w"), ("pp_k", "pp_k")):
            den = np.where(o[ok] != 0, np.abs(o[ok]), 1.0)
            assert np.max(np.abs(getattr(g, gk) - o[ok]) / den) <= tol, gk
        scale = np.maximum(np.abs(o["pp_gt"]), 1e-300)
        # per patch pair the |terms| scale is not available: use the per-bin scale
        for gk, ok, ak in (("pp_gt", "pp_gt", "gt_abs"), ("pp_gx", "pp_gx", "gx_abs")):
            assert np.max(np.abs(getattr(g, gk) - o[ok]) / np.maximum(o[ak], 1e-300)) <= tol, gk
        del scale
    if marks:
        # marks come back as (n_bin, n_A) views; compare against the oracle's
        # (n_A, n_bin).  W and K marks: relative to the lens' row; the signed
        # xi marks: relative to a bound of their sum of |terms|, the oracle's
        # kappa marks of |val| against |s1| + |s2| (|e1 c + e2 s| <= |e1| + |e2|)
        s1, s2 = ("e1", "e2") if kind == SHEAR else ("kappa_E", "kappa_B")
        o_abs = orc.correlate(KAPPA, dict(x=lens["x"], y=lens["y"], val=np.abs(lens["val"]), w=lens["w"]),
                              dict(x=src["x"], y=src["y"], kappa_E=np.abs(src[s1]) + np.abs(src[s2]),
                                   kappa_B=np.zeros(len(src["x"])), w=src["w"]), bins.bin_edges, marks=True)
        for gk, ok in (("xi_E_marks", "xi_e"), ("xi_B_marks", "xi_b"), ("K_marks", "k_marks"),
                       ("w_marks", "w_marks")):
            gm = np.asarray(getattr(g, gk), np.float64).T
            om = o[ok]
            if ok in ("xi_e", "xi_b"):
                scale = np.abs(o_abs["xi_e"]) + 1e-300
            else:
                scale = np.abs(om).sum(axis=1, keepdims=True) + 1e-300
            err = np.max(np.abs(gm - om) / scale)
            assert err <= 10 * tol, (gk, err)
        np.testing.assert_allclose(g.marks_weight, o["marks_weight"], rtol=10 * tol, atol=1e-300)
    return g, o


@pytest.mark.parametrize("kind", [KAPPA, SHEAR])
def test_uniform_log_bins(tp, orc, ctx, kind):
    lens, src = uniform_case(11 + kind, 4000, 5000, 100.0, kind)
    bins = tp.create_bins(0.5, 30.0, 15, tp.LOG)
    check(tp, orc, kind, lens, src, bins, ctx=ctx)


@pytest.mark.parametrize("kind", [KAPPA, SHEAR])
def test_patches_and_marks(tp, orc, ctx, kind):
    lens, src = uniform_case(21 + kind, 3000, 3500, 80.0, kind)
    r = _rng(5)
    pa = r.integers(0, 4, len(lens["x"])).astype(np.int32)
    pb = (np.floor(src["x"] / 80.0 * 3).clip(0, 2)).astype(np.int32)
    bins = tp.create_bins(1.0, 25.0, 12, tp.LOG)
    check(tp, orc, kind, lens, src, bins, marks=True, patch_A=pa, patch_B=pb, n_pa=4, n_pb=3, ctx=ctx)


def test_kappa_auto_same_catalogue(tp, orc, ctx):
    r = _rng(31)
    n = 6000
    x, y = r.uniform(0, 120, n).astype(np.float32), r.uniform(0, 120, n).astype(np.float32)
    k = r.normal(0, 0.02, n).astype(np.float32)
    w = r.uniform(0.5, 1.5, n).astype(np.float32)
    lens = dict(x=x, y=y, val=k, w=w)
    src = dict(x=x, y=y, kappa_E=k, kappa_B=r.normal(0, 0.02, n).astype(np.float32), w=w)
    bins = tp.create_bins(1.0, 40.0, 20, tp.LOG)
    g, o = check(tp, orc, KAPPA, lens, src, bins, ctx=ctx)
    assert g.n_pairs.sum() % 2 == 0   # ordered pairs are counted twice (src/tpcf.f90:266, :486)


@pytest.mark.parametrize("spacing", [0, 2, 3])
def test_other_spacings(tp, orc, ctx, spacing):
    lens, src = uniform_case(41 + spacing, 3000, 3000, 60.0, KAPPA)
    bins = tp.create_bins(0.0 if spacing != 2 else 0.5, 20.0, 10, spacing, 0.7)
    check(tp, orc, KAPPA, lens, src, bins, ctx=ctx)


def test_shear_clustered(tp, orc, ctx):
    synth = __import__("importlib").import_module("2pcf_amd.synth")
    wl = synth.make_workload("t", 30000, 30000, 300.0, 0, 0.5, 30.0, 20, 6, seed=77, clustered=True, npix=256,
                             fwhm=10.0)
    bins = tp.create_bins(wl.x_min, wl.x_max, wl.n_bin, tp.LOG)
    check(tp, orc, SHEAR, wl.lens, wl.src, bins, patch_A=wl.patch_A, patch_B=wl.patch_B, n_pa=6, n_pb=6,
          ctx=ctx)


def test_shear_masks_and_zero_weights(tp, orc, ctx):
    lens, src = uniform_case(51, 3000, 4000, 70.0, SHEAR, z=True, zero_w=True)
    bins = tp.create_bins(0.5, 20.0, 12, tp.LOG)
    check(tp, orc, SHEAR, lens, src, bins, marks=True, ctx=ctx)


@pytest.mark.parametrize("kind", [KAPPA, SHEAR])
def test_tiling_leaves_results_unchanged(tp, orc, kind):
    """The source cell side (TPCF_OPT_CELL_SIZE) and the work-item band
    height (TPCF_OPT_SPLIT) only change how the pairs are visited: which
    segment visits fall into the fast, edge and multi classes, and how a
    chunk's rows are cut into items.  Counts stay bit-exact against the oracle
    and the sums within the bar, with zero weights, patches and marks (no z
    masks: a segment whose z range may fail the z test is always swept as a
    multi visit), from cells of 1/50 x_max to cells as wide as x_max: the
    class counters (TPCF_OPT_COUNTERS) show the examined pairs moving from
    mostly edge visits (one bin edge in the visit's distance range) to mostly
    multi visits."""
    lens, src = uniform_case(61 + kind, 6000, 8000, 60.0, kind, zero_w=True)
    bins = tp.create_bins(1.0, 40.0, 5, tp.LOG)
    r = _rng(9)
    pa = r.integers(0, 4, len(lens["x"])).astype(np.int32)
    pb = r.integers(0, 3, len(src["x"])).astype(np.int32)
    c = tp.Context()
    try:
        c.set_option(2, 1.0)
        shares = {}
        for cell, split in ((0.75, 0.2), (3.0, 0.1), (12.0, 0.5), (40.0, 1.0), (0.0, 0.2)):
            c.set_option(1, cell)
            c.set_option(4, split)
            g, _ = check(tp, orc, kind, lens, src, bins, marks=True, patch_A=pa, patch_B=pb, n_pa=4, n_pb=3, ctx=c)
            pv = np.asarray(g.stats["pair_visits"], np.float64)
            shares[cell] = pv / pv.sum()
            print(f"cell {cell} split {split}: fast / edge / multi shares of the examined pairs {shares[cell]}")
        assert shares[0.75][1] > 0.5 and shares[40.0][2] > 0.5, shares
        assert shares[40.0][2] > shares[0.75][2] + 0.3, shares
    finally:
        c.close()


def test_automatic_cell_size(tp):
    """The automatic source cell side is sqrt(t * area / n_B) over the
    sources' extent (each side at least x_max / 64), clamped to
    [x_max / 512, 4 x_max], with t = 128 mean sources per cell for FP32
    geometry and t = 64 for FP64; TPCF_OPT_CELL_TARGET sets t for both."""
    lens, src = uniform_case(71, 3000, 20000, 100.0, SHEAR)
    src["y"] = (src["y"] * np.float32(0.3)).astype(np.float32)   # a 100' x 30' field
    bins = tp.create_bins(1.0, 10.0, 5, tp.LOG)

    def expected(t):
        x, y = src["x"].astype(np.float64), src["y"].astype(np.float64)
        fl = bins.x_max / 64.0
        area = max(np.ptp(x), fl) * max(np.ptp(y), fl)
        return min(max(np.sqrt(t * area / len(x)), bins.x_max / 512.0), 4.0 * bins.x_max)

    c = tp.Context()
    try:
        for target in (None, 32.0):
            if target is not None:
                c.set_option(3, target)
            for precision, t in (("f32", 128.0), ("f64", 64.0)):
                g = tp.correlate(SHEAR, lens, src, bins, precision=precision, real_bytes=4, ctx=c)
                want = expected(target or t)
                print(f"{precision} target {target}: cell {g.stats['cell_size']:.6f}, expected {want:.6f}")
                assert g.stats["cell_size"] == pytest.approx(want, rel=1e-12)
    finally:
        c.close()


@pytest.mark.parametrize("seed", range(32))
def test_random_configurations(tp, orc, ctx, seed):
    """Seeded random cases against the oracle at the bars above: catalogue
    sizes from 1 to a few thousand (lens chunks of 1, 63, 64 and 65 lenses,
    single sources), field sides of 5' to 120', 1 to 80 bins of every
    spacing (more than 63 bins: one pass per group of 63), x_min = 0 where the
    spacing allows it, both kinds, FP32 and FP64 inputs, patches, marks, z
    masks and zero weights."""
    r = _rng(1000 + seed)
    kind = int(r.integers(0, 2))
    n_a = int(r.choice([1, 7, 63, 64, 65, 200, 1500, 3000]))
    n_b = int(r.choice([1, 5, 64, 300, 2000, 4000]))
    side = float(r.uniform(5.0, 120.0))
    lens, src = uniform_case(2000 + seed, n_a, n_b, side, kind, z=bool(r.random() < 0.5),
                             zero_w=bool(r.random() < 0.5))
    precision = "f64" if r.random() < 0.3 else "f32"
    if precision == "f64":
        lens = {k: v.astype(np.float64) for k, v in lens.items()}
        src = {k: v.astype(np.float64) for k, v in src.items()}
    spacing = int(r.integers(0, 4))
    n_bin = int(r.choice([1, 3, 12, 40, 63, 64, 80]))
    x_min = 0.0 if spacing in (0, 3) and r.random() < 0.5 else float(r.uniform(0.1, 1.5))
    x_max = float(r.uniform(x_min + 1.0, max(x_min + 2.0, 0.7 * side)))
    bins = tp.create_bins(x_min, x_max, n_bin, spacing, 0.7)
    kw = {}
    if r.random() < 0.5:
        n_pa, n_pb = int(r.integers(1, 6)), int(r.integers(1, 6))
        kw = dict(patch_A=r.integers(0, n_pa, n_a).astype(np.int32), patch_B=r.integers(0, n_pb, n_b).astype(np.int32),
                  n_pa=n_pa, n_pb=n_pb)
    marks = bool(r.random() < 0.5)
    print(f"\nseed {seed}: kind {kind}, {n_a} x {n_b} on {side:.1f}', {precision}, spacing {spacing}, "
          f"{n_bin} bins [{x_min:.3f}, {x_max:.3f}], patches {bool(kw)}, marks {marks}", flush=True)
    check(tp, orc, kind, lens, src, bins, precision=precision, marks=marks, ctx=ctx, **kw)


def test_shear_marks_m_zero_cancelling_weights(tp, orc, ctx):
    """m == 0 (the K marks are copied from the W marks after the kernel) and
    lenses whose first bin holds only two sources of weights +w and -w: the
    bin's W and K marks are exactly 0 while xi_E / xi_B are not, so a marks
    pass that skipped lanes by their weight sum alone would lose them."""
    lens, src = uniform_case(52, 3000, 4000, 70.0, SHEAR)
    src["m"][:] = 0
    bins = tp.create_bins(1.0, 20.0, 12, tp.LOG)
    e1 = bins.bin_edges[1]
    planted = np.arange(0, 3000, 300)
    keep = np.ones(len(src["x"]), bool)
    for i in planted:
        keep &= np.hypot(src["x"] - lens["x"][i], src["y"] - lens["y"][i]) > 1.2 * e1
    src = {k: v[keep] for k, v in src.items()}
    r = _rng(53)
    add = {k: [] for k in src}
    for i in planted:
        ang = r.uniform(0, 2 * np.pi)
        for sgn in (1.0, -1.0):
            add["x"].append(lens["x"][i] + 1.1 * np.cos(ang))
            add["y"].append(lens["y"][i] + 1.1 * np.sin(ang))
            add["z"].append(1.0)
            add["e1"].append(r.normal(0, 0.27))
            add["e2"].append(r.normal(0, 0.27))
            add["w"].append(sgn * 0.75)
            add["m"].append(0.0)
    src = {k: np.concatenate([v, np.asarray(add[k], np.float32)]) for k, v in src.items()}
    g, o = check(tp, orc, SHEAR, lens, src, bins, marks=True, ctx=ctx)
    for i in planted:
        assert o["w_marks"][i, 0] == 0 and o["k_marks"][i, 0] == 0 and o["xi_e"][i, 0] != 0
        assert g.w_marks[0, i] == 0 and g.K_marks[0, i] == 0
        np.testing.assert_allclose(g.xi_E_marks[0, i], o["xi_e"][i, 0], rtol=1e-6)
        np.testing.assert_allclose(g.xi_B_marks[0, i], o["xi_b"][i, 0], rtol=1e-6)


@pytest.mark.parametrize("kind", [KAPPA, SHEAR])
@pytest.mark.parametrize("n_bin,spacing,x_min", [(63, 1, 0.3), (1, 1, 2.0), (63, 0, 0.0), (64, 1, 0.3)])
def test_bin_count_extremes(tp, orc, ctx, kind, n_bin, spacing, x_min):
    """The largest bin count of one pair-kernel pass (63: every lane of the
    per-bin table but one, the two-bin window at the last bin), one more (two
    passes, the second of a single bin), and a single bin, with marks and
    patches."""
    lens, src = uniform_case(91 + n_bin + spacing + kind, 2500, 3000, 60.0, kind, z=kind == SHEAR, zero_w=True)
    r = _rng(92)
    pa = r.integers(0, 3, len(lens["x"])).astype(np.int32)
    pb = r.integers(0, 2, len(src["x"])).astype(np.int32)
    bins = tp.create_bins(x_min, 25.0, n_bin, spacing)
    g, o = check(tp, orc, kind, lens, src, bins, marks=True, patch_A=pa, patch_B=pb, n_pa=3, n_pb=2, ctx=ctx)
    assert np.count_nonzero(g.n_pairs) == n_bin


@pytest.mark.parametrize("kind", [KAPPA, SHEAR])
@pytest.mark.parametrize("real_bytes", [4, 8])
def test_device_marks(tp, ctx, kind, real_bytes):
    """TPCF_DEVICE_MARKS: the marks written into device arrays equal the
    marks copied to host arrays (the FP64 atomics add in any order: equal to
    1e-12 of each lens' row scale), lenses outside the pair set (zero weight,
    or no source in range) included; and an empty pair set zeroes them."""
    import torch
    lens, src = uniform_case(58 + kind, 2500, 3000, 70.0, kind, zero_w=True)
    lens["x"][:50] += 500.0   # far from every source
    bins = tp.create_bins(0.5, 20.0, 12, tp.LOG)
    h = tp.correlate(kind, lens, src, bins, marks=True, real_bytes=real_bytes, ctx=ctx)
    d = tp.correlate(kind, lens, src, bins, marks=True, real_bytes=real_bytes, ctx=ctx, marks_device=True)
    np.testing.assert_array_equal(d.n_pairs, h.n_pairs)
    for k in ("xi_E_marks", "xi_B_marks", "K_marks", "w_marks"):
        dv, hv = getattr(d, k), np.asarray(getattr(h, k), np.float64)
        assert isinstance(dv, torch.Tensor) and dv.is_cuda and tuple(dv.shape) == hv.shape
        dv = dv.cpu().numpy().astype(np.float64)
        scale = np.abs(hv).sum(axis=0, keepdims=True) + 1e-300
        assert np.max(np.abs(dv - hv) / scale) <= (1e-6 if real_bytes == 4 else 1e-12), k
        assert np.all(dv[:, :50] == 0), k
    np.testing.assert_allclose(d.marks_weight.cpu().numpy(), h.marks_weight, rtol=1e-6 if real_bytes == 4 else 1e-12)
    far = {k: (v + 1000.0 if k == "x" else v) for k, v in src.items()}
    e = tp.correlate(kind, lens, far, bins, marks=True, real_bytes=real_bytes, ctx=ctx, marks_device=True)
    assert e.n_pairs.sum() == 0
    for k in ("xi_E_marks", "xi_B_marks", "K_marks", "w_marks", "marks_weight"):
        assert not torch.any(getattr(e, k) != 0), k


@pytest.mark.parametrize("kind", [KAPPA, SHEAR])
def test_fp64_path(tp, orc, ctx, kind):
    lens, src = uniform_case(61 + kind, 3000, 3000, 60.0, kind)
    lens = {k: v.astype(np.float64) + (1e-9 if k in ("x", "y") else 0) for k, v in lens.items()}
    src = {k: v.astype(np.float64) for k, v in src.items()}
    bins = tp.create_bins(0.5, 20.0, 12, tp.LOG)
    check(tp, orc, kind, lens, src, bins, precision="f64", marks=True, ctx=ctx)


@pytest.mark.parametrize("kind", [KAPPA, SHEAR])
def test_exact_edge_hits(tp, orc, ctx, kind):
    # integer lattices: many separations are exactly on integer bin edges and
    # on x_max, so FP32 d^2 equals the edge and the FP64 re-test decides
    g1 = np.arange(0, 24, dtype=np.float32)
    X, Y = np.meshgrid(g1, g1)
    x, y = X.ravel(), Y.ravel()
    n = len(x)
    r = _rng(71)
    lens = dict(x=x, y=y, val=r.normal(0, 0.02, n).astype(np.float32), w=r.uniform(0.5, 1.5, n).astype(np.float32),
                z=np.zeros(n, np.float32))
    if kind == KAPPA:
        src = dict(x=x, y=y, kappa_E=lens["val"], kappa_B=r.normal(0, 0.02, n).astype(np.float32), w=lens["w"])
        bins = tp.create_bins(0.0, 10.0, 10, tp.LIN)
    else:
        src = dict(x=x, y=y, z=np.ones(n, np.float32), e1=r.normal(0, 0.27, n).astype(np.float32),
                   e2=r.normal(0, 0.27, n).astype(np.float32), w=r.uniform(0.5, 1.5, n).astype(np.float32),
                   m=np.zeros(n, np.float32))
        bins = tp.create_bins(1.0, 10.0, 9, tp.LIN)
    g, o = check(tp, orc, kind, lens, src, bins, ctx=ctx)
    # d = 5 (3-4-5 triangles) sits on the edge at 5: the pair belongs to the bin above
    assert g.n_pairs[-1] > 0


def test_power_of_two_log_edges(tp, orc, ctx):
    g1 = np.arange(0, 40, dtype=np.float32)
    X, Y = np.meshgrid(g1, g1)
    x, y = X.ravel(), Y.ravel()
    n = len(x)
    lens = dict(x=x, y=y, val=np.ones(n, np.float32), w=np.ones(n, np.float32))
    src = dict(x=x, y=y, kappa_E=np.ones(n, np.float32), kappa_B=np.zeros(n, np.float32), w=np.ones(n, np.float32))
    bins = tp.create_bins(1.0, 32.0, 5, tp.LOG)   # edges 1, 2, 4, 8, 16, 32 exactly
    g, o = check(tp, orc, KAPPA, lens, src, bins, ctx=ctx)
    np.testing.assert_allclose(g.g_t[g.n_pairs > 0], 1.0, rtol=1e-12)   # constant kappa field: xi = 1


def test_empty_inputs(tp, ctx):
    bins = tp.create_bins(1.0, 10.0, 5, tp.LOG)
    lens = dict(x=np.zeros(0, np.float32), y=np.zeros(0, np.float32), val=np.zeros(0, np.float32),
                w=np.zeros(0, np.float32))
    src = dict(x=np.ones(3, np.float32), y=np.ones(3, np.float32), kappa_E=np.ones(3, np.float32),
               kappa_B=np.ones(3, np.float32), w=np.ones(3, np.float32))
    g = tp.correlate(KAPPA, lens, src, bins, ctx=ctx)
    assert g.n_pairs.sum() == 0 and np.all(np.isnan(g.mean_theta)) and np.all(g.g_t == 0)


@pytest.mark.parametrize("marks_device", [False, True])
def test_all_lenses_masked_with_marks(tp, ctx, marks_device):
    """Every lens weight 0 (no lens is used, src/tpcf.f90:285): no pairs, NaN
    mean_theta, g = 0, and the marks outputs and marks_weight all zero --
    written over whatever the output arrays held."""
    lens, src = uniform_case(95, 500, 700, 30.0, SHEAR)
    lens["w"][:] = 0
    bins = tp.create_bins(1.0, 10.0, 6, tp.LOG)
    g = tp.correlate(SHEAR, lens, src, bins, marks=True, ctx=ctx, marks_device=marks_device)
    assert g.n_pairs.sum() == 0 and np.all(np.isnan(g.mean_theta)) and np.all(g.g_t == 0)
    for k in ("xi_E_marks", "xi_B_marks", "K_marks", "w_marks", "marks_weight"):
        v = getattr(g, k)
        v = v.cpu().numpy() if marks_device else np.asarray(v)
        assert np.all(v == 0), k


def test_error_paths(tp, ctx):
    bins = tp.create_bins(1.0, 10.0, 5, tp.LOG)
    lens, src = uniform_case(81, 100, 100, 20.0, KAPPA)
    lens["x"][3] = np.nan
    with pytest.raises(tp.TpcfError) as e:
        tp.correlate(KAPPA, lens, src, bins, ctx=ctx)
    assert e.value.code == 6
    lens, src = uniform_case(82, 100, 100, 20.0, KAPPA)
    with pytest.raises(tp.TpcfError) as e:
        tp.correlate(KAPPA, lens, src, bins, patch_A=np.full(100, 5, np.int32), n_patch_A=4, ctx=ctx)
    assert e.value.code == 8
    field = tp.LensingSurveyField(lenses=tp.LensCatalog(**lens),
                                  convergence=tp.KappaMap(**src))
    many = tp.create_bins(1.0, 10.0, 64, tp.LOG)
    with pytest.raises(tp.TpcfError) as e:   # the adjoint takes at most 63 bins
        tp.adjoint(KAPPA, lens, src, many, np.ones(64), np.zeros(64), ctx=ctx)
    assert e.value.code == 5
    # the context stays usable after errors
    g = tp.kappa_2pcf(field, bins, None, False, 1, True, False)
    assert g.n_pairs.sum() > 0


def test_linearity_and_weight_scaling(tp, ctx):
    lens, src = uniform_case(91, 20000, 20000, 200.0, SHEAR)
    bins = tp.create_bins(0.5, 40.0, 16, tp.LOG)
    a = tp.correlate(SHEAR, lens, src, bins, ctx=ctx, real_bytes=8)
    lens2 = dict(lens, val=lens["val"] * 2)
    b = tp.correlate(SHEAR, lens2, src, bins, ctx=ctx, real_bytes=8)
    np.testing.assert_array_equal(a.n_pairs, b.n_pairs)
    # FP64 sums reached through atomics in no fixed order: equal to FP64
    # rounding, measured against the largest bin (a bin's sum can cancel)
    for k in ("sum_gt", "sum_gx"):
        ga, gb = getattr(a, k), getattr(b, k)
        np.testing.assert_allclose(gb, 2 * ga, rtol=1e-12, atol=1e-13 * np.abs(ga).max())
    src3 = dict(src, w=src["w"] * 4)
    c = tp.correlate(SHEAR, lens, src3, bins, ctx=ctx, real_bytes=8)
    np.testing.assert_allclose(c.g_t, a.g_t, rtol=1e-9)
    np.testing.assert_allclose(c.sum_w, 4 * a.sum_w, rtol=1e-12)


def test_zero_separation_with_x_min_zero(tp, orc, ctx):
    # coincident lens/source pairs with x_min = 0 land in bin 0 with an
    # undefined position angle: the reference's g_t/g_x of bin 0 are NaN
    r = _rng(101)
    n = 500
    x = r.uniform(0, 30, n).astype(np.float32)
    y = r.uniform(0, 30, n).astype(np.float32)
    lens = dict(x=x, y=y, val=r.normal(0, 0.02, n).astype(np.float32), w=r.uniform(0.5, 1.5, n).astype(np.float32),
                z=np.zeros(n, np.float32))
    src = dict(x=x.copy(), y=y.copy(), z=np.ones(n, np.float32), e1=r.normal(0, 0.27, n).astype(np.float32),
               e2=r.normal(0, 0.27, n).astype(np.float32), w=r.uniform(0.5, 1.5, n).astype(np.float32),
               m=np.zeros(n, np.float32))
    bins = tp.create_bins(0.0, 5.0, 5, tp.LIN)
    g = tp.correlate(SHEAR, lens, src, bins, ctx=ctx, real_bytes=8)
    o = orc.correlate(SHEAR, lens, src, bins.bin_edges)
    np.testing.assert_array_equal(g.n_pairs, o["n_pairs"])
    assert np.isnan(o["gt"][0]) and np.isnan(g.sum_gt[0]) and np.isnan(g.sum_gx[0])
    assert np.all(np.isfinite(g.sum_gt[1:]))
    np.testing.assert_allclose(g.sum_w, o["w"], rtol=1e-6)
    np.testing.assert_allclose(g.sum_gt[1:], o["gt"][1:], rtol=0, atol=1e-6 * o["gt_abs"][1:].max())
