"""The oracle's block bootstrap and covariance (oracle.bootstrap, oracle.covariance).

Pinned by
  * the reference's own covariance_matrix (src/utils.f90:745-764, FP32, built
    from its sources into oracle/_ref) on correlated samples: within 2e-6 of
    the largest entry; a 1/n normalisation or a missing mean fails that bar;
  * the resampling's defining cases, with block sums taken from an oracle
    correlation with patches (itself pinned against the reference):
    unit multiplicities give the full-sample g_t, g_x of the correlation;
    one block drawn alone gives that block's own G/K; doubling every
    multiplicity leaves g unchanged and doubles the weights; a block-bin with
    K = 0 everywhere gives 0 (src/tpcf.f90:1183-1189).
"""
import numpy as np
import pytest

SHEAR = 0


def test_covariance_matches_reference(orc):
    if not orc.have_reference():
        pytest.skip("reference build (oracle/_ref) not available")
    r = np.random.default_rng(1)
    x = r.normal(0, 1, (500, 12)) * np.linspace(1, 3, 12) + r.normal(0, 1, (500, 1)) + np.linspace(2, 5, 12)
    c, rc = orc.covariance(x), orc.ref_covariance(x)
    scale = np.max(np.abs(c))
    assert np.max(np.abs(c - rc)) <= 2e-6 * scale
    assert np.max(np.abs(c * (499 / 500) - rc)) > 1e-3 * scale           # 1/n instead of 1/(n-1)
    raw = x.T @ x / 499
    assert np.max(np.abs(raw - rc)) > 1e-2 * scale                       # means not subtracted


def _blocks(orc):
    r = np.random.default_rng(7)
    n_a, n_b, P = 3000, 3000, 9
    lens = dict(x=r.uniform(0, 60, n_a), y=r.uniform(0, 60, n_a), z=np.zeros(n_a), val=r.normal(0, 0.02, n_a),
                w=r.uniform(0.5, 1.5, n_a))
    src = dict(x=r.uniform(0, 60, n_b), y=r.uniform(0, 60, n_b), z=np.ones(n_b), e1=r.normal(0, 0.27, n_b),
               e2=r.normal(0, 0.27, n_b), w=r.uniform(0.5, 1.5, n_b), m=r.normal(0, 0.05, n_b))
    pa = (np.floor(lens["x"] / 20) + 3 * np.floor(lens["y"] / 20)).astype(np.int32)
    e = orc.edges(1, np.float32(0.5), np.float32(15.0), 8)
    o = orc.correlate(SHEAR, lens, src, e, pa, np.zeros(n_b, np.int32), P, 1, patches=True)
    blocks = dict(g_t=o["pp_gt"].sum(axis=0), g_x=o["pp_gx"].sum(axis=0), k=o["pp_k"].sum(axis=0),
                  weight=np.bincount(pa, weights=lens["w"], minlength=P))
    return o, blocks, P


def test_bootstrap_defining_cases(orc):
    o, blocks, P = _blocks(orc)
    # unit multiplicities: the measurement itself
    b = orc.bootstrap(blocks, np.ones((1, P), dtype=np.int32))
    np.testing.assert_allclose(b["g_t"][0], o["g_t"], rtol=1e-12, atol=1e-18)
    np.testing.assert_allclose(b["g_x"][0], o["g_x"], rtol=1e-12, atol=1e-18)
    assert np.isclose(b["weights"][0], blocks["weight"].sum())
    # one block alone, and doubled multiplicities
    M = np.zeros((P, P), dtype=np.int32)
    M[np.arange(P), np.arange(P)] = 3
    b = orc.bootstrap(blocks, M)
    np.testing.assert_allclose(b["g_t"], blocks["g_t"] / blocks["k"], rtol=1e-12)
    r = np.random.default_rng(3)
    M = r.integers(0, 4, (20, P)).astype(np.int32)
    b1, b2 = orc.bootstrap(blocks, M), orc.bootstrap(blocks, 2 * M)
    np.testing.assert_allclose(b1["g_t"], b2["g_t"], rtol=1e-12)
    np.testing.assert_allclose(2 * b1["weights"], b2["weights"], rtol=1e-12)
    # a 0.1 % error in K is caught by the unit case
    bad = dict(blocks, k=blocks["k"] * 1.001)
    assert not np.allclose(orc.bootstrap(bad, np.ones((1, P), np.int32))["g_t"][0], o["g_t"], rtol=1e-6)
    # empty block-bin: 0
    z = dict(blocks, k=blocks["k"].copy(), g_t=blocks["g_t"].copy())
    z["k"][:, 0] = 0
    z["g_t"][:, 0] = 0
    assert np.all(orc.bootstrap(z, M)["g_t"][:, 0] == 0)
