"""The CPU oracle pinned against the reference's own code and against
closed forms (CPU only).

  * golden vectors written by the reference's create_bins / find_bin_index and
    kd-tree search (tests/golden/ref_*.npz, from oracle/_ref/ref_harness);
  * closed forms: exact edges, exact lattice separations, pure tangential and
    cross shear patterns, constant fields;
  * invariants: brute force == pruned visit, linearity, masks == removal,
    per-patch sums == per-bin sums == summed marks;
  * the FP32 kd-tree port (CPU baseline) against the oracle.
"""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
SHEAR, KAPPA = 0, 1


def near_edge(x, edges, rel=1e-5):
    e = np.asarray(edges, np.float64)[None, :]
    return np.any(np.abs(np.asarray(x, np.float64)[:, None] - e) <= rel * np.maximum(e, 1e-30), axis=1)


# ------------------------------------------------------------- binning -----
def test_edges_closed_forms(orc):
    np.testing.assert_array_equal(orc.edges(orc.LOG, 1.0, 1024.0, 10), 2.0 ** np.arange(11))
    np.testing.assert_array_equal(orc.edges(orc.LOG, 0.25, 64.0, 8), 0.25 * 2.0 ** np.arange(9))
    np.testing.assert_array_equal(orc.edges(orc.LIN, 0.0, 10.0, 10), np.arange(11.0))
    np.testing.assert_array_equal(orc.edges(orc.SQRT, 0.0, 100.0, 10), np.arange(11.0) ** 2)
    e = orc.edges(orc.IHS, 0.5, 20.0, 12, 0.7)
    p = float(np.float32(0.7))          # ihs_theta is a single-precision argument
    t = np.arcsinh(p * e) / p
    np.testing.assert_allclose(np.diff(t), (t[-1] - t[0]) / 12, rtol=1e-12)
    assert e[0] == np.float32(0.5) and e[-1] == np.float32(20.0)


def test_bin_membership_definition(orc):
    e = orc.edges(orc.LOG, 1.0, 1024.0, 10)
    d = np.array([0.999, 1.0, 1.5, 2.0, 1023.9, 1024.0, 1024.0001, 0.0])
    np.testing.assert_array_equal(orc.bin_of_d2(d * d, e), [-1, 0, 0, 1, 9, 9, -1, -1])


def test_binning_matches_reference_find_bin_index(orc):
    g = np.load(os.path.join(GOLD, "ref_bins.npz"))
    for c in range(int(g["n_cases"])):
        sp, lo, hi, nb, p = g[f"case{c}_params"]
        sp, nb = int(sp), int(nb)
        x = g[f"case{c}_x"].astype(np.float64)
        ref = g[f"case{c}_idx"].astype(np.int64)            # 1-based, -1 outside
        e = orc.edges(sp, float(lo), float(hi), nb, float(p))
        ours = orc.bin_of_d2(x * x, e).astype(np.int64)     # 0-based, -1 outside
        ours = np.where(ours >= 0, ours + 1, -1)
        x_in = x[x >= 0]
        mism = (ours != ref)[x >= 0]
        # the reference decides in FP32 through log10/sqrt/asinh: the two
        # definitions agree except within 1e-6 (relative) of an edge
        assert not np.any(mism & ~near_edge(x_in, e, 1e-6)), (c, x_in[mism & ~near_edge(x_in, e, 1e-6)][:5])
        # and exactly on the random part of the sample
        assert not np.any(mism[:4000])
        # the FP32 port used for CPU timing restates find_bin_index: identical
        port = orc.rp_find_bin(x.astype(np.float32), float(lo), float(hi), nb, sp, float(p)).astype(np.int64)
        np.testing.assert_array_equal(port, ref)
        # create_bins: FP32 reference edges vs FP64 closed form
        np.testing.assert_allclose(g[f"case{c}_edges"][1:-1], e[1:-1], rtol=2e-6)
        assert g[f"case{c}_edges"][-1] == np.float32(hi)


def test_bin_centers_match_reference(orc, tp):
    g = np.load(os.path.join(GOLD, "ref_bins.npz"))
    for c in range(int(g["n_cases"])):
        sp, lo, hi, nb, p = g[f"case{c}_params"]
        b = tp.create_bins(float(lo), float(hi), int(nb), int(sp), float(p))
        np.testing.assert_allclose(b.bin_centers, g[f"case{c}_centers"], rtol=2e-6, atol=1e-7)


def test_pair_set_matches_reference_kdtree(orc):
    g = np.load(os.path.join(GOLD, "ref_kdtree.npz"))
    px, py, qx, qy = (g[k].astype(np.float64) for k in ("px", "py", "qx", "qy"))
    r = float(g["r"][0])
    counts = g["counts"]
    # the reference's own property (src/test_kdtree.f90): tree search == brute force
    assert np.array_equal(counts[:, 0], counts[:, 1])
    np.testing.assert_array_equal(g["tree_idx"], g["brute_idx"])
    # the oracle's pair set: FP64 d^2 <= x_max^2 on the FP32 coordinates
    pos = 0
    for i in range(len(qx)):
        k = counts[i, 0]
        ref = set((g["tree_idx"][pos:pos + k] - 1).tolist())
        pos += k
        dx, dy = qx[i] - px, qy[i] - py
        d2 = dx * dx + dy * dy
        ours = set(np.nonzero(d2 <= r * r)[0].tolist())
        diff = ref ^ ours
        assert all(abs(d2[j] - r * r) <= 1e-5 * r * r for j in diff)


# ------------------------------------------------------------ correlation ----
def uniform(seed, n_a, n_b, side, kind, z=False):
    r = np.random.default_rng(seed)
    lens = dict(x=r.uniform(0, side, n_a).astype(np.float32), y=r.uniform(0, side, n_a).astype(np.float32),
                val=r.normal(0, 0.02, n_a), w=r.uniform(0.5, 1.5, n_a),
                z=r.uniform(-0.2, 1.0, n_a) if z else np.zeros(n_a))
    if kind == SHEAR:
        src = dict(x=r.uniform(0, side, n_b).astype(np.float32), y=r.uniform(0, side, n_b).astype(np.float32),
                   z=r.uniform(-0.2, 1.2, n_b) if z else np.ones(n_b), e1=r.normal(0, .27, n_b),
                   e2=r.normal(0, .27, n_b), w=r.uniform(0.5, 1.5, n_b), m=r.normal(0, .02, n_b))
    else:
        src = dict(x=r.uniform(0, side, n_b).astype(np.float32), y=r.uniform(0, side, n_b).astype(np.float32),
                   kappa_E=r.normal(0, .02, n_b), kappa_B=r.normal(0, .02, n_b), w=r.uniform(0.5, 1.5, n_b))
    return lens, src


@pytest.mark.parametrize("kind", [SHEAR, KAPPA])
def test_brute_equals_pruned(orc, kind):
    lens, src = uniform(3 + kind, 1500, 1800, 60.0, kind, z=(kind == SHEAR))
    e = orc.edges(orc.LOG, 0.5, 15.0, 10)
    a = orc.correlate(kind, lens, src, e, brute=True, marks=True)
    b = orc.correlate(kind, lens, src, e, brute=False, marks=True)
    np.testing.assert_array_equal(a["n_pairs"], b["n_pairs"])
    for k in ("theta", "gt", "gx", "w", "k"):
        np.testing.assert_allclose(a[k], b[k], rtol=1e-13, atol=1e-300)
    np.testing.assert_allclose(a["xi_e"], b["xi_e"], rtol=1e-12, atol=1e-18)


def test_lattice_counts_exact(orc):
    # integer lattice: d^2 is an exact integer, bins [k, k+1) in d, last bin closed
    g = np.arange(0, 15, dtype=np.float64)
    X, Y = np.meshgrid(g, g)
    x, y = X.ravel(), Y.ravel()
    n = len(x)
    lens = dict(x=x, y=y, val=np.ones(n), w=np.ones(n))
    src = dict(x=x, y=y, kappa_E=np.ones(n), kappa_B=np.zeros(n), w=np.ones(n))
    e = orc.edges(orc.LIN, 0.0, 8.0, 8)
    res = orc.correlate(KAPPA, lens, src, e, brute=True)
    d2 = ((x[:, None] - x[None, :]) ** 2 + (y[:, None] - y[None, :]) ** 2).astype(np.int64).ravel()
    want = np.zeros(8, np.int64)
    for k in range(8):
        lo, hi = k * k, (k + 1) * (k + 1)
        want[k] = np.count_nonzero((d2 >= lo) & (d2 < hi)) if k < 7 else np.count_nonzero((d2 >= lo) & (d2 <= hi))
    np.testing.assert_array_equal(res["n_pairs"], want)
    np.testing.assert_allclose(res["g_t"], 1.0, rtol=1e-14)            # constant field
    np.testing.assert_allclose(res["w"], want.astype(float), rtol=0)


def ring_sources(radii, n_per, gamma, cross=False):
    """Sources on circles around a lens at the origin with ellipticity of a
    pure tangential (or, rotated by 45 degrees, pure cross) pattern.  With
    dx = x_A - x_B the position angle phi of (dx, dy) equals alpha + pi."""
    xs, ys, e1, e2 = [], [], [], []
    for r in radii:
        a = np.linspace(0, 2 * np.pi, n_per, endpoint=False) + 0.1
        xs.append(r * np.cos(a)); ys.append(r * np.sin(a))
        if not cross:   # gamma_t = -(e1 cos2phi + e2 sin2phi) = gamma
            e1.append(-gamma * np.cos(2 * a)); e2.append(-gamma * np.sin(2 * a))
        else:           # gamma_x = -(-e1 sin2phi + e2 cos2phi) = gamma
            e1.append(gamma * np.sin(2 * a)); e2.append(-gamma * np.cos(2 * a))
    n = n_per * len(radii)
    return dict(x=np.concatenate(xs), y=np.concatenate(ys), z=np.ones(n), e1=np.concatenate(e1),
                e2=np.concatenate(e2), w=np.ones(n), m=np.zeros(n))


@pytest.mark.parametrize("cross", [False, True])
def test_pure_tangential_and_cross_patterns(orc, cross):
    lens = dict(x=np.zeros(1), y=np.zeros(1), z=np.zeros(1), val=np.ones(1), w=np.ones(1))
    e = orc.edges(orc.LOG, 1.0, 64.0, 6)
    src = ring_sources([1.5, 3.0, 6.0, 12.0, 24.0, 48.0], 37, 0.1, cross)
    res = orc.correlate(SHEAR, lens, src, e, brute=True)
    assert np.all(res["n_pairs"] == 37)
    big, zero = ("g_x", "g_t") if cross else ("g_t", "g_x")
    np.testing.assert_allclose(res[big], 0.1, rtol=1e-12)
    np.testing.assert_allclose(res[zero], 0.0, atol=1e-14)
    np.testing.assert_allclose(res["mean_theta"], [1.5, 3.0, 6.0, 12.0, 24.0, 48.0], rtol=1e-14)


def test_multiplicative_bias_enters_k(orc):
    lens = dict(x=np.zeros(1), y=np.zeros(1), z=np.zeros(1), val=np.ones(1), w=np.ones(1))
    e = orc.edges(orc.LOG, 1.0, 64.0, 6)
    src = ring_sources([1.5, 3.0, 6.0, 12.0, 24.0, 48.0], 20, 0.1)
    src["m"] = np.full(len(src["x"]), 0.25)
    res = orc.correlate(SHEAR, lens, src, e, brute=True)
    # K = sum w (1 + m): g_t = gamma / (1 + m)
    np.testing.assert_allclose(res["g_t"], 0.1 / 1.25, rtol=1e-12)


def test_masks_equal_removal(orc):
    lens, src = uniform(7, 800, 900, 40.0, SHEAR, z=True)
    e = orc.edges(orc.LOG, 0.5, 12.0, 8)
    full = orc.correlate(SHEAR, lens, src, e, brute=True)
    keep_a = (lens["z"] >= 0)
    keep_b = (src["z"] >= 0)
    red = orc.correlate(SHEAR, {k: v[keep_a] for k, v in lens.items()}, {k: v[keep_b] for k, v in src.items()}, e,
                        brute=True)
    np.testing.assert_array_equal(full["n_pairs"], red["n_pairs"])
    # with z_A > z_B excluded, setting all source z above all lens z can only add pairs
    src2 = dict(src, z=np.where(src["z"] >= 0, 2.0, src["z"]))
    more = orc.correlate(SHEAR, lens, src2, e, brute=True)
    assert np.all(more["n_pairs"] >= full["n_pairs"])


def test_zero_weights(orc):
    lens, src = uniform(8, 600, 700, 40.0, KAPPA)
    e = orc.edges(orc.LOG, 0.5, 12.0, 8)
    lens["w"][:100] = 0.0      # skipped lenses (src/tpcf.f90:474)
    src["w"][:100] = 0.0       # kept sources: counted with zero weight
    full = orc.correlate(KAPPA, lens, src, e, brute=True, marks=True)
    red = orc.correlate(KAPPA, {k: v[100:] for k, v in lens.items()}, src, e, brute=True)
    np.testing.assert_array_equal(full["n_pairs"], red["n_pairs"])
    assert np.all(full["w_marks"][:100] == 0) and np.all(full["marks_weight"][:100] == 0)
    red2 = orc.correlate(KAPPA, {k: v[100:] for k, v in lens.items()}, {k: v[100:] for k, v in src.items()}, e,
                         brute=True)
    assert np.all(full["n_pairs"] >= red2["n_pairs"])
    np.testing.assert_allclose(full["w"], red2["w"], rtol=1e-13)


def test_linearity_and_weight_invariance(orc):
    lens, src = uniform(9, 700, 800, 40.0, SHEAR)
    e = orc.edges(orc.LOG, 0.5, 12.0, 8)
    a = orc.correlate(SHEAR, lens, src, e)
    b = orc.correlate(SHEAR, dict(lens, val=3 * lens["val"]), src, e)
    np.testing.assert_allclose(b["gt"], 3 * a["gt"], rtol=1e-13)
    c = orc.correlate(SHEAR, lens, dict(src, w=7 * src["w"]), e)
    np.testing.assert_allclose(c["g_t"], a["g_t"], rtol=1e-12)
    np.testing.assert_allclose(c["mean_theta"], a["mean_theta"], rtol=1e-12)


def test_kappa_auto_symmetry(orc):
    r = np.random.default_rng(10)
    n = 900
    x, y = r.uniform(0, 50, n), r.uniform(0, 50, n)
    k1, k2, w = r.normal(0, .02, n), r.normal(0, .02, n), r.uniform(.5, 1.5, n)
    e = orc.edges(orc.LOG, 0.5, 15.0, 9)
    ab = orc.correlate(KAPPA, dict(x=x, y=y, val=k1, w=w), dict(x=x, y=y, kappa_E=k2, kappa_B=k2, w=w), e)
    ba = orc.correlate(KAPPA, dict(x=x, y=y, val=k2, w=w), dict(x=x, y=y, kappa_E=k1, kappa_B=k1, w=w), e)
    np.testing.assert_array_equal(ab["n_pairs"], ba["n_pairs"])
    np.testing.assert_allclose(ab["gt"], ba["gt"], rtol=1e-12)
    assert np.all(ab["n_pairs"] % 2 == 0)   # ordered pairs counted twice


def test_patch_sums_and_marks(orc):
    lens, src = uniform(11, 700, 800, 40.0, SHEAR)
    e = orc.edges(orc.LOG, 0.5, 12.0, 8)
    r = np.random.default_rng(4)
    pa, pb = r.integers(0, 3, 700).astype(np.int32), r.integers(0, 4, 800).astype(np.int32)
    res = orc.correlate(SHEAR, lens, src, e, pa, pb, 3, 4, marks=True, patches=True)
    np.testing.assert_array_equal(res["pp_n"].sum(axis=(0, 1)), res["n_pairs"])
    for q, k in (("pp_gt", "gt"), ("pp_w", "w"), ("pp_k", "k"), ("pp_theta", "theta")):
        np.testing.assert_allclose(res[q].sum(axis=(0, 1)), res[k], rtol=1e-12)
    # summed over source patches = per-lens-patch block sums of the marks (src/tpcf.f90:1130-1137)
    for p in range(3):
        np.testing.assert_allclose(res["pp_gt"][:, p, :].sum(axis=0), res["xi_e"][pa == p].sum(axis=0), rtol=1e-11,
                                   atol=1e-15)
        np.testing.assert_allclose(res["pp_k"][:, p, :].sum(axis=0), res["k_marks"][pa == p].sum(axis=0), rtol=1e-11)
    np.testing.assert_allclose(res["marks_weight"], res["w_marks"].sum(axis=1), rtol=1e-13)


def test_empty_bin_gives_nan_mean_theta(orc):
    lens = dict(x=np.zeros(1), y=np.zeros(1), val=np.ones(1), w=np.ones(1))
    src = dict(x=np.array([3.0]), y=np.zeros(1), kappa_E=np.ones(1), kappa_B=np.ones(1), w=np.ones(1))
    res = orc.correlate(KAPPA, lens, src, orc.edges(orc.LOG, 1.0, 16.0, 4), brute=True)
    assert res["n_pairs"].tolist() == [0, 1, 0, 0]
    assert np.isnan(res["mean_theta"][[0, 2, 3]]).all() and res["mean_theta"][1] == 3.0
    assert res["g_t"][0] == 0.0   # K > 0 guard (src/tpcf.f90:371-376)


@pytest.mark.parametrize("kind", [SHEAR, KAPPA])
def test_reference_port_agrees_with_oracle(orc, kind):
    lens, src = uniform(12 + kind, 2000, 2500, 80.0, kind)
    e = orc.edges(orc.LOG, 0.5, 20.0, 10)
    o = orc.correlate(kind, lens, src, e)
    for tree in (True, False):
        p = orc.ref_port(kind, lens, src, 0.5, 20.0, 10, orc.LOG, n_threads=4, use_tree=tree, marks=True)
        # FP32 log10 binning may move a pair sitting on an edge by one bin
        assert np.abs(p["n_pairs"] - o["n_pairs"]).max() <= 3
        np.testing.assert_allclose(p["weights"], o["w"], rtol=2e-4)
        np.testing.assert_allclose(p["g_t"], o["g_t"], rtol=5e-3, atol=1e-3 * np.abs(o["g_t"]).max())
        np.testing.assert_allclose(p["marks_weight"].sum(), o["w"].sum(), rtol=1e-4)


# ------------------------------------------ the reference's own correlation ----
import refcases  # noqa: E402


def oracle_on_case(orc, c):
    kind, sp, lo, hi, nb, p, lens, src, ref = refcases.load(c)
    e = orc.edges(sp, lo, hi, nb, p)
    o = orc.correlate(kind, lens, src, e, marks=True)
    got = dict(n_pairs=o["n_pairs"], weights=o["w"], g_t=o["g_t"], g_x=o["g_x"], mean_theta=o["mean_theta"],
               xi_e=o["xi_e"], xi_b=o["xi_b"], k_marks=o["k_marks"], w_marks=o["w_marks"],
               marks_weight=o["marks_weight"])
    return kind, lens, src, e, ref, got


@pytest.mark.parametrize("c", range(refcases.n_cases()))
def test_oracle_matches_reference_correlation(orc, c):
    kind, lens, src, e, ref, got = oracle_on_case(orc, c)
    refcases.compare(kind, lens, src, e, ref, got)


def test_reference_comparison_catches_mistakes(orc):
    """A dropped term, a sign flip, a swapped component or a wrong mask in the
    oracle's output must fail the comparison with the reference."""
    kind, lens, src, e, ref, got = oracle_on_case(orc, 0)
    scan = refcases.scan_pairs(kind, lens, src, e)
    refcases.compare(kind, lens, src, e, ref, got, scan)
    band = scan[0]
    k = int(np.argmax(np.where(band.sum(axis=0) == 0, got["n_pairs"], -1)))     # a bin clear of the edge band
    i = int(np.argmax(np.where(band.sum(axis=1) == 0, np.abs(got["xi_e"]).sum(axis=1), -1)))  # a lens likewise
    bad = [dict(got, g_x=-got["g_x"]), dict(got, g_t=got["g_x"], g_x=got["g_t"]),
           dict(got, n_pairs=got["n_pairs"] + np.eye(len(e) - 1, dtype=np.int64)[k]),
           dict(got, g_t=got["g_t"] * 1.02), dict(got, weights=got["weights"] * (1 + 1e-3)),
           dict(got, xi_e=got["xi_e"] * np.where(np.arange(got["xi_e"].shape[0])[:, None] == i, 1.01, 1.0))]
    # the z_A > z_B cut left out
    src2 = dict(src, z=np.full(len(src["z"]), 5.0, np.float32) * (src["z"] >= 0) - (src["z"] < 0))
    o = orc.correlate(kind, lens, src2, e)
    bad.append(dict(got, n_pairs=o["n_pairs"]))
    for b in bad:
        with pytest.raises(AssertionError):
            refcases.compare(kind, lens, src, e, ref, b, scan)
