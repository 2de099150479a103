"""The oracle's spherical geometry (x, y = longitude, latitude in radians).

Pinned by
  * the reference's own distance_sphere / angles_sphere (src/utils.f90:512-528,
    :618-639; FP32, built from its sources into oracle/_ref) on random pairs
    at separations where FP32 resolves them (0.3-2.8 rad: angle within 5e-6
    relative, cos 2phi and sin 2phi within 2e-5), with planted mistakes (a
    sign of the north component, swapped factors) failing the same bars;
  * the reference's own tangential_shear_tpcf / scalar_kappa_tpcf with
    spherical_coords (src/tpcf.f90:190-545, kd-tree and brute force) on
    catalogues around longitude pi (where its spherical search runs without
    rotating the field): counts equal except pairs within 1e-3 (relative) of
    an edge, where the reference's FP32 acos decides differently; g_t, g_x
    within 1e-3 of the bin's sum of |terms| / K (the FP32 acos of a 2-degree
    separation is good to 1e-4 relative), weights within 3e-4;
  * closed forms: separations along the equator and along a meridian are the
    coordinate differences; a source due east or west of the lens has
    cos 2phi = 1, due north or south -1, and north-east (small scale)
    sin 2phi = 1;
  * the flat-sky limit: a 0.01 x 0.01 rad patch on the equator gives the flat
    correlation of (x, y) to 1e-4 of each sum's scale.
"""
import numpy as np
import pytest

SHEAR, KAPPA = 0, 1


def need_ref(orc):
    if not orc.have_reference():
        pytest.skip("reference build (oracle/_ref) not available")


def test_sph_pair_matches_reference_functions(orc):
    need_ref(orc)
    r = np.random.default_rng(4)
    n = 20000
    lon1, lon2 = r.uniform(0, 2 * np.pi, n), r.uniform(0, 2 * np.pi, n)
    lat1, lat2 = r.uniform(-1.2, 1.2, n), r.uniform(-1.2, 1.2, n)
    f32 = lambda v: v.astype(np.float32).astype(np.float64)   # noqa: E731
    d, c2, s2 = orc.sph_pair(f32(lon1), f32(lat1), f32(lon2), f32(lat2))
    rd, rc2, rs2 = orc.ref_sph_pair(lon1, lat1, lon2, lat2)
    m = (d > 0.3) & (d < np.pi - 0.3)
    assert m.sum() > 15000
    assert np.max(np.abs(rd - d)[m] / d[m]) < 5e-6
    assert np.max(np.abs(rc2 - c2)[m]) < 2e-5
    assert np.max(np.abs(rs2 - s2)[m]) < 2e-5
    # the same bars reject plausible mistakes
    assert np.max(np.abs(rc2 + c2)[m]) > 0.1            # wrong sign of cos 2phi
    assert np.max(np.abs(rs2 + s2)[m]) > 0.1            # north component negated
    assert np.max(np.abs(rc2 - s2)[m]) > 0.1            # factors swapped


def test_sph_closed_forms(orc):
    lon0, lat0 = 1.0, 0.0
    h = 1e-3
    # east, west, north, south, north-east
    lon2 = np.array([lon0 + h, lon0 - h, lon0, lon0, lon0 + 1e-5])
    lat2 = np.array([lat0, lat0, lat0 + h, lat0 - h, lat0 + 1e-5])
    d, c2, s2 = orc.sph_pair(np.full(5, lon0), np.full(5, lat0), lon2, lat2)
    np.testing.assert_allclose(d[:4], h, rtol=1e-12)
    np.testing.assert_allclose(c2[:4], [1, 1, -1, -1], atol=1e-12)
    np.testing.assert_allclose(s2[:4], 0, atol=1e-12)
    np.testing.assert_allclose(s2[4], 1.0, atol=1e-9)
    # separations along a meridian at any longitude, and along the equator
    lat = np.linspace(-1.4, 1.3, 7)
    d, _, _ = orc.sph_pair(np.full(7, 2.0), lat, np.full(7, 2.0), lat + 0.2)
    np.testing.assert_allclose(d, 0.2, rtol=1e-12)
    d, _, _ = orc.sph_pair(np.zeros(3), np.zeros(3), np.array([0.5, 2.0, 3.0]), np.zeros(3))
    np.testing.assert_allclose(d, [0.5, 2.0, 3.0], rtol=1e-12)
    # the squared chord edges of the bin test
    e = np.array([0.0, 0.01, 1.0, np.pi, 4.0])
    np.testing.assert_allclose(orc.chord_edges_squared(e), (2 * np.sin(np.minimum(e, np.pi) / 2)) ** 2, rtol=1e-15)


def _sph_catalogues(seed, n_a, n_b, lon_c, half, masks=True):
    r = np.random.default_rng(seed)
    f = np.float32
    lens = dict(x=r.uniform(lon_c - half, lon_c + half, n_a).astype(f), y=r.uniform(-half, half, n_a).astype(f),
                val=r.normal(0, 0.02, n_a).astype(f), w=r.uniform(0.5, 1.5, n_a).astype(f))
    src = dict(x=r.uniform(lon_c - half, lon_c + half, n_b).astype(f), y=r.uniform(-half, half, n_b).astype(f),
               w=r.uniform(0.5, 1.5, n_b).astype(f))
    lens["z"] = (r.uniform(-0.1, 1.0, n_a) if masks else np.zeros(n_a)).astype(f)
    src["z"] = (r.uniform(-0.1, 1.2, n_b) if masks else np.ones(n_b)).astype(f)
    src["e1"] = r.normal(0, 0.27, n_b).astype(f)
    src["e2"] = r.normal(0, 0.27, n_b).astype(f)
    src["m"] = (r.normal(0, 0.05, n_b) if masks else np.zeros(n_b)).astype(f)
    src["kappa_E"] = r.normal(0, 0.02, n_b).astype(f)
    src["kappa_B"] = r.normal(0, 0.02, n_b).astype(f)
    if masks:
        lens["w"][::23] = 0
    return lens, src


def _near_edge_pairs(lens, src, edges, band):
    """Per bin, the pairs whose angle lies within `band` (relative) of one of
    the bin's edges (all pairs, masks ignored: an upper bound)."""
    from numpy import cos, sin
    ua = np.stack([cos(lens["y"].astype(float)) * cos(lens["x"].astype(float)),
                   cos(lens["y"].astype(float)) * sin(lens["x"].astype(float)), sin(lens["y"].astype(float))], 1)
    ub = np.stack([cos(src["y"].astype(float)) * cos(src["x"].astype(float)),
                   cos(src["y"].astype(float)) * sin(src["x"].astype(float)), sin(src["y"].astype(float))], 1)
    d = np.arccos(np.clip(ua @ ub.T, -1, 1)).ravel()
    out = np.zeros(len(edges) - 1, dtype=np.int64)
    for k in range(len(edges) - 1):
        for e in (edges[k], edges[k + 1]):
            out[k] += np.count_nonzero(np.abs(d - e) <= band * e)
    return out


@pytest.mark.parametrize("kind", [SHEAR, KAPPA])
def test_sph_correlation_matches_reference(orc, kind):
    need_ref(orc)
    lens, src = _sph_catalogues(11 + kind, 1500, 2000, np.pi, 0.3)
    if kind == KAPPA:
        lens.pop("z")
    xmin, xmax, nbin = np.radians(2.0), np.radians(20.0), 8
    e = orc.edges(1, np.float32(xmin), np.float32(xmax), nbin)
    o = orc.correlate(kind, lens, src, e, spherical=True)
    amb = _near_edge_pairs(lens, src, e, 1e-3)
    for tree in (True, False):
        ref = orc.reference(kind, lens, src, xmin, xmax, nbin, 1, spherical=True, use_tree=tree)
        assert np.all(np.abs(ref["n_pairs"] - o["n_pairs"]) <= amb), (ref["n_pairs"], o["n_pairs"], amb)
        assert np.all(o["n_pairs"] > 1000)
        scale_t = o["gt_abs"] / o["k"]
        scale_x = o["gx_abs"] / o["k"]
        assert np.all(np.abs(ref["g_t"] - o["g_t"]) <= 1e-3 * scale_t)
        assert np.all(np.abs(ref["g_x"] - o["g_x"]) <= 1e-3 * scale_x)
        np.testing.assert_allclose(ref["weights"], o["w"], rtol=3e-4)
        np.testing.assert_allclose(ref["mean_theta"], o["mean_theta"], rtol=1e-4)


@pytest.mark.parametrize("kind", [SHEAR, KAPPA])
def test_sph_flat_limit(orc, kind):
    """A 0.01-rad patch on the equator: the sphere is flat to 1e-5."""
    r = np.random.default_rng(21)
    n_a, n_b = 1500, 1500
    lens = dict(x=r.uniform(0.5, 0.51, n_a), y=r.uniform(-0.005, 0.005, n_a), z=np.zeros(n_a),
                val=r.normal(0, 0.02, n_a), w=r.uniform(0.5, 1.5, n_a))
    src = dict(x=r.uniform(0.5, 0.51, n_b), y=r.uniform(-0.005, 0.005, n_b), z=np.ones(n_b),
               e1=r.normal(0, 0.27, n_b), e2=r.normal(0, 0.27, n_b), w=r.uniform(0.5, 1.5, n_b), m=np.zeros(n_b),
               kappa_E=r.normal(0, 0.02, n_b), kappa_B=r.normal(0, 0.02, n_b))
    e = orc.edges(1, np.float32(2e-4), np.float32(5e-3), 6)
    s = orc.correlate(kind, lens, src, e, spherical=True, brute=True)
    f = orc.correlate(kind, lens, src, e, brute=True)
    # counts: only pairs within ~1e-5 of an edge may move
    assert np.all(np.abs(s["n_pairs"] - f["n_pairs"]) <= 1e-3 * f["n_pairs"] + 2)
    for key, scale in (("gt", f["gt_abs"]), ("gx", f["gx_abs"]), ("w", f["w"]), ("theta", f["theta"])):
        assert np.all(np.abs(s[key] - f[key]) <= 1e-4 * scale + 2e-3 * scale / np.sqrt(f["n_pairs"])), key
