"""Spherical coordinates (TPCF_SPHERICAL) on the GPU against the FP64 oracle.

x, y are longitude and latitude in radians, the bins angles in radians.  The
GPU decides bins on squared chords of FP64 unit vectors (device sincos), the
oracle on squared chords of its own FP64 unit vectors (host sincos), so a pair
may fall on the other side of an edge only when its squared chord lies within
1e-9 (relative) of the edge's: counts are compared exactly after allowing for
the pairs in that band (counted here with numpy), which are none in these
catalogues unless the data put one there.  Sums: FP64 path, within 1e-10 of the
bin's scale (Theta, W, K relative; G_t, G_x against the sum of |terms|).
Cases: a field around longitude pi with z masks, m, zero lens weights,
patches and marks (FP64 inputs); a cap around the north pole that also
crosses longitude 0 (FP32 inputs, so the FP32 entry point widens them); a
kappa-kappa auto-correlation across longitude 0; bins from 5 to 150 degrees
over the whole sphere (separations above 18 degrees take the half-angle path
of the angle, not the series alone); x_min = 0 with every object also on the
other side (zero separations: angle 0, and NaN position angles for shear as
in the reference); the Python shear_2pcf wrapper with
spherical_coords (the spherical adjoint is tested in
test_gpu_adjoint_sphere.py).
Seeded random fields anywhere on the sphere, of every size down to a single
object, with log or linear bins out to up to 60 degrees, close the list.
The first three cases' bins span the separation of about 4.2 degrees
(squared chord 5.4e-3) below which a segment visit takes the short angle
series, so visits of both variants are compared.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
SHEAR, KAPPA = 0, 1


def unit(lon, lat):
    lon, lat = np.asarray(lon, np.float64), np.asarray(lat, np.float64)
    return np.stack([np.cos(lat) * np.cos(lon), np.cos(lat) * np.sin(lon), np.sin(lat)], 1)


def edge_band_pairs(lens, src, e2c, band=1e-9):
    """Per bin: pairs whose squared chord is within `band` of either edge (an
    edge at 0 is not counted: a zero chord is in bin 0 on both sides)."""
    ua, ub = unit(lens["x"], lens["y"]), unit(src["x"], src["y"])
    out = np.zeros(len(e2c) - 1, dtype=np.int64)
    for i0 in range(0, len(ua), 512):
        d = ua[i0:i0 + 512, None, :] - ub[None, :, :]
        c2 = np.einsum("ijk,ijk->ij", d, d).ravel()
        for k in range(len(e2c) - 1):
            for e in (e2c[k], e2c[k + 1]):
                if e > 0:
                    out[k] += np.count_nonzero(np.abs(c2 - e) <= band * e)
    return out


def field(seed, n_a, n_b, kind, lon, lat, masks=True, dtype=np.float64):
    r = np.random.default_rng(seed)
    lens = dict(x=lon(r, n_a), y=lat(r, n_a), val=r.normal(0, 0.02, n_a), w=r.uniform(0.5, 1.5, n_a))
    if masks:
        lens["w"][::19] = 0.0
    if kind == SHEAR:
        lens["z"] = r.uniform(-0.1, 1.0, n_a) if masks else np.zeros(n_a)
        src = dict(x=lon(r, n_b), y=lat(r, n_b), z=r.uniform(-0.1, 1.2, n_b) if masks else np.ones(n_b),
                   e1=r.normal(0, 0.27, n_b), e2=r.normal(0, 0.27, n_b), w=r.uniform(0.5, 1.5, n_b),
                   m=r.normal(0, 0.05, n_b) if masks else np.zeros(n_b))
    else:
        src = dict(x=lon(r, n_b), y=lat(r, n_b), kappa_E=r.normal(0, 0.02, n_b), kappa_B=r.normal(0, 0.02, n_b),
                   w=r.uniform(0.5, 1.5, n_b))
    cast = lambda d: {k: np.asarray(v, dtype) for k, v in d.items()}   # noqa: E731
    return cast(lens), cast(src)


def check(tp, orc, ctx, kind, lens, src, x_min, x_max, n_bin, precision, patches=0, marks=False, tol=1e-10,
          spacing=None, min_pairs=10000):
    bins = tp.create_bins(x_min, x_max, n_bin, tp.LOG if spacing is None else spacing)
    kw = {}
    pa = pb = None
    if patches:
        r = np.random.default_rng(99)
        pa = r.integers(0, patches, len(lens["x"])).astype(np.int32)
        pb = r.integers(0, patches, len(src["x"])).astype(np.int32)
        kw = dict(patch_A=pa, patch_B=pb, n_patch_A=patches, n_patch_B=patches)
    g = tp.correlate(kind, lens, src, bins, precision=precision, marks=marks, real_bytes=8, ctx=ctx, spherical=True,
                     **kw)
    o = orc.correlate(kind, lens, src, bins.bin_edges, pa, pb, max(patches, 1), max(patches, 1), marks=marks,
                      patches=bool(patches), spherical=True)
    amb = edge_band_pairs(lens, src, orc.chord_edges_squared(bins.bin_edges))
    print("\npairs per bin:", g.n_pairs.tolist(), "near-edge pairs:", amb.tolist())
    assert np.all(np.abs(g.n_pairs - o["n_pairs"]) <= amb)
    assert g.n_pairs.sum() >= min_pairs
    exact = amb.sum() == 0
    for gk, ok in (("sum_theta", "theta"), ("sum_w", "w"), ("sum_k", "k")):
        den = np.where(o[ok] != 0, np.abs(o[ok]), 1.0)
        err = np.max(np.abs(getattr(g, gk) - o[ok]) / den)
        print(gk, err)
        assert err <= (tol if exact else 1e-6), gk
    for gk, ok, ak in (("sum_gt", "gt", "gt_abs"), ("sum_gx", "gx", "gx_abs")):
        den = np.where(o[ak] != 0, o[ak], 1.0)
        err = np.max(np.abs(getattr(g, gk) - o[ok]) / den)
        print(gk, err)
        assert err <= (tol if exact else 1e-6), gk
    if patches and exact:
        np.testing.assert_array_equal(g.pp_n, o["pp_n"])
        for gk in ("pp_theta", "pp_w", "pp_k"):
            den = np.where(o[gk] != 0, np.abs(o[gk]), 1.0)
            assert np.max(np.abs(getattr(g, gk) - o[gk]) / den) <= tol, gk
        for gk, ak in (("pp_gt", "gt_abs"), ("pp_gx", "gx_abs")):
            assert np.max(np.abs(getattr(g, gk) - o[gk]) / np.maximum(o[ak], 1e-300)) <= tol, gk
    if marks and exact:
        for gk, ok in (("xi_E_marks", "xi_e"), ("xi_B_marks", "xi_b"), ("K_marks", "k_marks"), ("w_marks", "w_marks")):
            a = np.asarray(getattr(g, gk)).T
            scale = np.maximum(np.abs(o[ok]).max(axis=1, keepdims=True), 1e-300)
            assert np.max(np.abs(a - o[ok]) / scale) <= 1e-9, gk
    return g, o


def test_sphere_shear_masks_patches_marks(tp, orc, ctx):
    lens, src = field(31, 2500, 3000, SHEAR, lambda r, n: r.uniform(np.pi - 0.15, np.pi + 0.15, n),
                      lambda r, n: r.uniform(-0.15, 0.15, n))
    check(tp, orc, ctx, SHEAR, lens, src, np.radians(0.5), np.radians(8.0), 10, "f64", patches=4, marks=True)


def test_sphere_pole_and_meridian_f32_inputs(tp, orc, ctx):
    # a cap of 20 degrees around the north pole: every longitude, including 0 / 2 pi
    lens, src = field(32, 2500, 3000, SHEAR, lambda r, n: r.uniform(0, 2 * np.pi, n),
                      lambda r, n: np.arcsin(r.uniform(np.sin(np.radians(70)), 1.0, n)), dtype=np.float32)
    check(tp, orc, ctx, SHEAR, lens, src, np.radians(0.5), np.radians(10.0), 10, "f32")


def test_sphere_kappa_auto_across_meridian(tp, orc, ctx):
    r = np.random.default_rng(33)
    n = 3000
    lon = np.mod(r.uniform(-0.12, 0.12, n), 2 * np.pi)
    lat = r.uniform(-0.12, 0.12, n)
    kap, w = r.normal(0, 0.02, n), r.uniform(0.5, 1.5, n)
    lens = dict(x=lon, y=lat, val=kap, w=w)
    src = dict(x=lon, y=lat, kappa_E=kap, kappa_B=r.normal(0, 0.02, n), w=w)
    check(tp, orc, ctx, KAPPA, lens, src, np.radians(0.3), np.radians(6.0), 12, "f64")


def test_sphere_wide_angles_all_sky(tp, orc, ctx):
    # uniform over the sphere: most pairs lie between 18 and 150 degrees
    lens, src = field(35, 1200, 1500, SHEAR, lambda r, n: r.uniform(0, 2 * np.pi, n),
                      lambda r, n: np.arcsin(r.uniform(-1.0, 1.0, n)))
    g, o = check(tp, orc, ctx, SHEAR, lens, src, np.radians(5.0), np.radians(150.0), 8, "f64", patches=3,
                 marks=True)
    # the mean angle of each bin lies inside the bin
    edges = np.asarray(tp.create_bins(np.radians(5.0), np.radians(150.0), 8, tp.LOG).bin_edges)
    mt = g.sum_theta / g.sum_w
    assert np.all((mt > edges[:-1]) & (mt < edges[1:]))


def test_sphere_zero_separation(tp, orc, ctx):
    """x_min = 0 on linear bins with the same positions on both sides, across
    longitude 0: every object's self-pair sits at zero separation in bin 0.
    Kappa-kappa: the self-pairs' angle is 0 and every sum matches the oracle.
    Shear: their position angle is 0/0, so G_t, G_x of bin 0 and every lens'
    xi_E / xi_B marks of bin 0 are NaN, as in the oracle (the reference
    divides by sin d = 0 in angles_sphere, src/utils.f90:618-639); Theta, W,
    K and the other bins match."""
    r = np.random.default_rng(36)
    n = 2000
    lon = np.mod(r.uniform(-0.06, 0.06, n), 2 * np.pi)
    lat = r.uniform(0.2, 0.32, n)
    kap, w = r.normal(0, 0.02, n), r.uniform(0.5, 1.5, n)
    x_max = np.radians(3.0)
    check(tp, orc, ctx, KAPPA, dict(x=lon, y=lat, val=kap, w=w),
          dict(x=lon, y=lat, kappa_E=kap, kappa_B=r.normal(0, 0.02, n), w=w), 0.0, x_max, 6, "f64",
          spacing=tp.LIN)
    lens = dict(x=lon, y=lat, z=np.zeros(n), val=kap, w=w)
    src = dict(x=lon, y=lat, z=np.ones(n), e1=r.normal(0, 0.27, n), e2=r.normal(0, 0.27, n),
               w=r.uniform(0.5, 1.5, n), m=np.zeros(n))
    bins = tp.create_bins(0.0, x_max, 6, tp.LIN)
    g = tp.correlate(SHEAR, lens, src, bins, precision="f64", marks=True, real_bytes=8, ctx=ctx, spherical=True)
    o = orc.correlate(SHEAR, lens, src, bins.bin_edges, None, None, 1, 1, marks=True, patches=False, spherical=True)
    amb = edge_band_pairs(lens, src, orc.chord_edges_squared(bins.bin_edges))
    print("\npairs per bin:", g.n_pairs.tolist(), "near-edge pairs:", amb.tolist())
    assert amb.sum() == 0
    np.testing.assert_array_equal(g.n_pairs, o["n_pairs"])
    assert g.n_pairs[0] >= n
    for gk, ok in (("sum_theta", "theta"), ("sum_w", "w"), ("sum_k", "k")):
        assert np.max(np.abs(getattr(g, gk) - o[ok]) / np.abs(o[ok])) <= 1e-10, gk
    for gk, ok, ak in (("sum_gt", "gt", "gt_abs"), ("sum_gx", "gx", "gx_abs")):
        have, want = getattr(g, gk), o[ok]
        assert np.isnan(have[0]) and np.isnan(want[0]), gk
        assert np.max(np.abs(have[1:] - want[1:]) / o[ak][1:]) <= 1e-10, gk
    for gk, ok in (("xi_E_marks", "xi_e"), ("xi_B_marks", "xi_b")):
        a = np.asarray(getattr(g, gk)).T
        assert np.isnan(a[:, 0]).all() and np.isnan(o[ok][:, 0]).all(), gk
        scale = np.maximum(np.abs(o[ok][:, 1:]).max(axis=1, keepdims=True), 1e-300)
        assert np.max(np.abs(a[:, 1:] - o[ok][:, 1:]) / scale) <= 1e-9, gk
    for gk, ok in (("K_marks", "k_marks"), ("w_marks", "w_marks")):
        a = np.asarray(getattr(g, gk)).T
        scale = np.maximum(np.abs(o[ok]).max(axis=1, keepdims=True), 1e-300)
        assert np.max(np.abs(a - o[ok]) / scale) <= 1e-9, gk


@pytest.mark.parametrize("seed", range(12))
def test_sphere_random_configurations(tp, orc, ctx, seed):
    """Seeded random spherical cases against the oracle: fields anywhere on
    the sphere (near the poles and across longitude 0 included), 1 to 3000
    objects, field half-widths of 0.02 to 0.4 rad, log or linear bins out to
    up to 60 degrees (the wide-angle path), both kinds, FP32 and FP64 inputs,
    patches, marks, masks."""
    r = np.random.default_rng(3000 + seed)
    kind = int(r.integers(0, 2))
    n_a = int(r.choice([1, 63, 65, 400, 1500, 3000]))
    n_b = int(r.choice([1, 64, 500, 2000, 3000]))
    lon0, lat0 = float(r.uniform(0, 2 * np.pi)), float(r.uniform(-1.5, 1.5))
    half = float(r.uniform(0.02, 0.4))

    def lon(rr, n):
        return np.mod(lon0 + rr.uniform(-half, half, n) / max(np.cos(lat0), 0.1), 2 * np.pi)

    def lat(rr, n):
        return np.clip(lat0 + rr.uniform(-half, half, n), -np.pi / 2, np.pi / 2)

    dtype = np.float32 if r.random() < 0.4 else np.float64
    lens, src = field(3100 + seed, n_a, n_b, kind, lon, lat, masks=bool(r.random() < 0.5), dtype=dtype)
    spacing = tp.LOG if r.random() < 0.7 else tp.LIN
    x_max = float(r.uniform(0.3, 1.0) * (np.radians(60.0) if r.random() < 0.2 else 2 * half))
    x_min = float(r.uniform(0.01, 0.2) * x_max)
    n_bin = int(r.choice([1, 5, 12, 30]))
    patches = int(r.choice([0, 3]))
    marks = bool(r.random() < 0.5)
    print(f"\nseed {seed}: kind {kind}, {n_a} x {n_b} around ({lon0:.2f}, {lat0:.2f}) half {half:.3f}, "
          f"{np.dtype(dtype).name}, {n_bin} bins [{x_min:.4f}, {x_max:.4f}] spacing {spacing}, patches {patches}, "
          f"marks {marks}", flush=True)
    check(tp, orc, ctx, kind, lens, src, x_min, x_max, n_bin, "f32" if dtype == np.float32 else "f64",
          patches=patches, marks=marks, spacing=spacing, min_pairs=0)


def test_sphere_python_wrapper(tp, orc):
    lens, src = field(34, 1500, 2000, SHEAR, lambda r, n: r.uniform(1.0, 1.2, n), lambda r, n: r.uniform(0.4, 0.6, n),
                      masks=False, dtype=np.float32)
    fld = tp.LensingSurveyField(lenses=tp.LensCatalog(x=lens["x"], y=lens["y"], z=lens["z"], val=lens["val"],
                                                      w=lens["w"]),
                                shapes=tp.ShapeCatalog(x=src["x"], y=src["y"], z=src["z"], e1=src["e1"],
                                                       e2=src["e2"], w=src["w"], m=src["m"]))
    bins = tp.create_bins(np.radians(0.5), np.radians(5.0), 8, tp.LOG)
    res = tp.shear_2pcf(fld, bins, spherical_coords=True)
    o = orc.correlate(SHEAR, lens, src, bins.bin_edges, spherical=True)
    np.testing.assert_array_equal(res.n_pairs, o["n_pairs"])
    np.testing.assert_allclose(res.mean_theta, o["mean_theta"], rtol=1e-6)
    scale = o["gt_abs"] / o["k"]
    assert np.all(np.abs(res.g_t - o["g_t"]) <= 1e-6 * scale)

