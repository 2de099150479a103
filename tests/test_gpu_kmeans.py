"""k-means patch labels on the GPU (tpcf_kmeans_*) against the oracle's
restatement of partition_data (src/kmeans.f90:12-263), itself pinned against
the reference's own code.

The GPU sums the centroids in 64-bit fixed point (integer atomics), so its
result does not depend on the order of the adds: two runs agree bit for bit.
Its arithmetic is not the oracle's, so a point at an exact tie between two
centroids could still go either way: labels must agree for >= 99.99 % of the
points (all of them in these cases so far), centroids within 1e-9 of the
field size, rounds within one.
Cases: clustered and uniform flat catalogues in FP64 and FP32, given starting
centroids, spherical coordinates across longitude 0, 10 seeded random
configurations; k outside [2, 1024] is refused.
The weight equaliser and the equal-size post-processing are compared with the
oracle's form that moves the cluster weights after every block of 256
consecutive points (block=256, the GPU's steps; the reference moves them
after every point, block=1, which the oracle also restates and pins against
the reference): weighted and unit-weight catalogues (ties
between cluster weights, where the reference's quicksort order of the
clusters matters), flat and spherical, with and without the equaliser, with
the same label, centroid and weight bounds; the post-processing's cluster
weights are also checked against the sums of the members' weights.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def blobs(seed, n, k=12, dtype=np.float64):
    r = np.random.default_rng(seed)
    cent = r.uniform(0, 100, (k, 2))
    lab = r.integers(0, k, n)
    x = cent[lab, 0] + r.normal(0, 4, n)
    y = cent[lab, 1] + r.normal(0, 4, n)
    return x.astype(dtype), y.astype(dtype), r.uniform(0.5, 1.5, n).astype(dtype)


def agree(g, o, size):
    frac = np.mean(g["labels"] == o["labels"])
    print(f"\nlabels agree {frac:.6f}, rounds {g['iterations']} / {o['iterations']}")
    assert frac >= 0.9999
    assert abs(g["iterations"] - o["iterations"]) <= 1
    m = np.isfinite(o["centroids"]).all(axis=1)
    assert np.max(np.abs(g["centroids"][m] - o["centroids"][m])) <= 1e-9 * size
    np.testing.assert_allclose(g["cluster_weights"][m], o["cluster_weights"][m], rtol=1e-9)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_kmeans_clustered(tp, orc, ctx, dtype):
    x, y, w = blobs(5, 200000, dtype=dtype)
    g = tp.kmeans(x, y, w, 64, ctx=ctx)
    o = orc.kmeans(x, y, w, 64)
    agree(g, o, 100.0)


def test_kmeans_uniform_given_centroids(tp, orc, ctx):
    r = np.random.default_rng(6)
    n = 300000
    x, y, w = r.uniform(0, 3000, n), r.uniform(0, 2000, n), r.uniform(0.5, 1.5, n)
    c0 = np.stack([r.uniform(0, 3000, 128), r.uniform(0, 2000, 128)], 1)
    g = tp.kmeans(x, y, w, 128, centroids=c0, max_iterations=40, ctx=ctx)
    o = orc.kmeans(x, y, w, 128, centroids=c0, max_iterations=40)
    agree(g, o, 3000.0)
    # reassignment tolerance stops early
    g2 = tp.kmeans(x, y, w, 128, centroids=c0, reassignment_tol=0.01, ctx=ctx)
    o2 = orc.kmeans(x, y, w, 128, centroids=c0, reassignment_tol=0.01)
    agree(g2, o2, 3000.0)
    assert g2["iterations"] < g["iterations"] or g["iterations"] == 41


def test_kmeans_spherical_across_meridian(tp, orc, ctx):
    r = np.random.default_rng(7)
    n = 100000
    lon = np.mod(r.uniform(-0.3, 0.3, n), 2 * np.pi)
    lat = r.uniform(-0.3, 0.3, n)
    w = r.uniform(0.5, 1.5, n)
    c0 = np.stack([np.mod(r.uniform(-0.3, 0.3, 16), 2 * np.pi), r.uniform(-0.3, 0.3, 16)], 1)
    g = tp.kmeans(lon, lat, w, 16, centroids=c0, spherical=True, max_iterations=30, ctx=ctx)
    o = orc.kmeans(lon, lat, w, 16, centroids=c0, spherical=True, max_iterations=30)
    agree(g, o, 2 * np.pi)


@pytest.mark.parametrize("seed", range(10))
def test_kmeans_random_configurations(tp, orc, ctx, seed):
    """Seeded random cases against the oracle at the bars above: 50 to 20000
    points, k from 2 to 64 (never above n), clustered or uniform fields,
    FP32 or FP64, the fixed starting grid or given centroids, with or without
    the weight equaliser (moved per step of 256 points, as on the GPU)."""
    r = np.random.default_rng(6000 + seed)
    n = int(r.choice([50, 500, 5000, 20000]))
    k = int(min(n, r.choice([2, 3, 17, 64])))
    dt = np.float32 if r.random() < 0.5 else np.float64
    if r.random() < 0.5:
        x, y, w = blobs(6100 + seed, n, k=int(r.integers(2, 20)), dtype=dt)
    else:
        x, y, w = (r.uniform(0, 100, n).astype(dt), r.uniform(0, 100, n).astype(dt),
                   r.uniform(0.5, 1.5, n).astype(dt))
    c0 = np.stack([r.uniform(0, 100, k), r.uniform(0, 100, k)], 1) if r.random() < 0.5 else None
    s = float(r.choice([0.0, 3.0]))
    print(f"\nseed {seed}: n {n}, k {k}, {np.dtype(dt).name}, given centroids {c0 is not None}, scaling {s}",
          flush=True)
    g = tp.kmeans(x, y, w, k, centroids=c0, weight_equalizer_scaling=s, max_iterations=50, ctx=ctx)
    o = orc.kmeans(x, y, w, k, centroids=c0, scaling=s, max_iterations=50, block=256 if s > 0 else None)
    agree(g, o, 100.0)


def test_kmeans_refusals(tp, ctx):
    x, y, w = blobs(8, 1000)
    with pytest.raises(tp.TpcfError):
        tp.kmeans(x, y, w, 1, ctx=ctx)
    with pytest.raises(tp.TpcfError):
        tp.kmeans(x, y, w, 2000, ctx=ctx)
    for bad in (dict(weight_equalizer_scaling=-1.0), dict(weight_equalizer_scaling=float("nan")),
                dict(equal_size=True, weight_tol=-0.1)):
        with pytest.raises(tp.TpcfError) as e:
            tp.kmeans(x, y, w, 8, ctx=ctx, **bad)
        assert e.value.code == tp._native.ERR_ARG, bad
    # a starting mode other than 0 (grid) or 1 (given centroids), through the C-ABI
    import ctypes as C
    N = tp._native
    o = N.KmeansOpts()
    o.k, o.max_iterations, o.init = 8, 5, 2
    lab, cen, cw = np.zeros(len(x), np.int32), np.zeros((8, 2)), np.zeros(8)
    rc = N.lib().tpcf_kmeans_f64(ctx.handle, len(x), x.ctypes.data, y.ctypes.data, w.ctypes.data, C.byref(o), 0,
                                 lab.ctypes.data, cen.ctypes.data, cw.ctypes.data, None, None)
    assert rc == N.ERR_ARG and "init" in ctx.last_error()
    # the context stays usable after the refusals
    assert tp.kmeans(x, y, w, 8, ctx=ctx)["iterations"] >= 1


def test_kmeans_reproducible(tp, ctx):
    """Two runs on the same input give identical labels, centroids and
    weights (no order-dependent floating-point adds)."""
    x, y, w = blobs(11, 300000, k=40)
    a = tp.kmeans(x, y, w, 64, ctx=ctx)
    b = tp.kmeans(x, y, w, 64, ctx=ctx)
    np.testing.assert_array_equal(a["labels"], b["labels"])
    np.testing.assert_array_equal(a["centroids"], b["centroids"])
    np.testing.assert_array_equal(a["cluster_weights"], b["cluster_weights"])
    assert a["iterations"] == b["iterations"]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_kmeans_weight_equaliser(tp, orc, ctx, dtype):
    x, y, w = blobs(21, 150000, k=20, dtype=dtype)
    g = tp.kmeans(x, y, w, 32, weight_equalizer_scaling=5.0, max_iterations=60, ctx=ctx)
    o = orc.kmeans(x, y, w, 32, scaling=5.0, max_iterations=60, block=256)
    agree(g, o, 100.0)
    plain = orc.kmeans(x, y, w, 32, max_iterations=60)
    assert np.mean(plain["labels"] == o["labels"]) < 0.99   # the equaliser changes the labels


@pytest.mark.parametrize("case", ["weighted", "unit_weights", "equalised", "sphere", "sphere_equalised"])
def test_kmeans_equal_size(tp, orc, ctx, case):
    r = np.random.default_rng(31)
    sph = case.startswith("sphere")
    if sph:
        n, k = 60000, 24
        x = r.uniform(1.0, 1.6, n)
        y = r.normal(0.3, 0.12, n).clip(-0.1, 0.7)
        w = r.uniform(0.5, 1.5, n)
        size = 2 * np.pi
    else:
        x, y, w = blobs(32, 120000, k=15, dtype=np.float32)
        k, size = 40, 100.0
        if case == "unit_weights":
            w = np.ones_like(w)
    s = 3.0 if case.endswith("equalised") else 0.0
    for tol in (0.0, 0.05):
        g = tp.kmeans(x, y, w, k, weight_equalizer_scaling=s, equal_size=True, weight_tol=tol, max_iterations=40,
                      spherical=sph, ctx=ctx)
        o = orc.kmeans(x, y, w, k, scaling=s, max_iterations=40, block=256, equal_size=True, weight_tol=tol,
                       spherical=sph)
        print(f"{case} tol {tol}: {o['pp_moves']} hand-overs", flush=True)
        assert o["pp_moves"] > 0
        agree(g, o, size)
        np.testing.assert_allclose(g["cluster_weights"],
                                   np.bincount(g["labels"], weights=np.asarray(w, np.float64), minlength=k),
                                   rtol=1e-9)
    again = tp.kmeans(x, y, w, k, weight_equalizer_scaling=s, equal_size=True, max_iterations=40, spherical=sph,
                      ctx=ctx)
    g0 = tp.kmeans(x, y, w, k, weight_equalizer_scaling=s, equal_size=True, max_iterations=40, spherical=sph,
                   ctx=ctx)
    np.testing.assert_array_equal(again["labels"], g0["labels"])
    np.testing.assert_array_equal(again["cluster_weights"], g0["cluster_weights"])
