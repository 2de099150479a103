"""The oracle's k-means (oracle.kmeans, orc_kmeans) against the reference's own
partition_data (src/kmeans.f90:12-263, FP32, one thread, fixed starting
centroids; built from its source into oracle/_ref).

* Voronoi form (no weight equaliser): the same labels for every point on
  clustered and on uniform catalogues, the same number of rounds, centroids
  within 1e-3 of the field size (the reference sums in FP32).
* Weight-equalised form (scaling 5, weights moved point by point): labels
  agree for >= 99 % of the points (FP32 against FP64 drift over 50 rounds).
* Planted mistakes fail: unweighted centroids, or moving the weights only at
  the end of a round (the GPU order), change the equalised labels for most
  points / measurably.
"""
import numpy as np
import pytest


def need_ref(orc):
    if not orc.have_reference():
        pytest.skip("reference build (oracle/_ref) not available")


def blobs(seed, n=20000, k=12):
    r = np.random.default_rng(seed)
    cent = r.uniform(0, 100, (k, 2))
    lab = r.integers(0, k, n)
    x = (cent[lab, 0] + r.normal(0, 4, n)).astype(np.float32)
    y = (cent[lab, 1] + r.normal(0, 4, n)).astype(np.float32)
    w = r.uniform(0.5, 1.5, n).astype(np.float32)
    return x, y, w


@pytest.mark.parametrize("case", ["blobs", "uniform"])
def test_kmeans_voronoi_matches_reference(orc, case):
    need_ref(orc)
    if case == "blobs":
        x, y, w = blobs(2)
    else:
        r = np.random.default_rng(3)
        x, y = r.uniform(0, 300, 15000).astype(np.float32), r.uniform(0, 200, 15000).astype(np.float32)
        w = r.uniform(0.5, 1.5, 15000).astype(np.float32)
    for k in (9, 16):
        ref = orc.ref_kmeans(x, y, w, k)
        o = orc.kmeans(x, y, w, k)
        assert np.mean(ref["labels"] == o["labels"]) >= (1.0 if case == "blobs" else 0.999)
        size = max(np.ptp(x), np.ptp(y))
        assert np.max(np.abs(ref["centroids"] - o["centroids"])) <= 1e-3 * size


def test_kmeans_equalised_matches_reference_and_mistakes_fail(orc):
    need_ref(orc)
    x, y, w = blobs(2)
    ref = orc.ref_kmeans(x, y, w, 16, scaling=5.0, max_iterations=50)
    o = orc.kmeans(x, y, w, 16, scaling=5.0, max_iterations=50)
    assert np.mean(ref["labels"] == o["labels"]) >= 0.99
    jac = orc.kmeans(x, y, w, 16, scaling=5.0, max_iterations=50, jacobi=True)
    assert np.mean(ref["labels"] == jac["labels"]) < 0.9
    unw = orc.kmeans(x, y, np.ones_like(w), 16)
    refv = orc.ref_kmeans(x, y, w, 16)
    assert np.max(np.abs(unw["centroids"] - refv["centroids"])) > 1e-3
