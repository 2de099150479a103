"""Block bootstrap and covariance on the GPU (tpcf_bootstrap) against the oracle.

* Given multiplicities: g_t, g_x equal the oracle's resampling
  (src/tpcf.f90:1165-1191) within 1e-12 of sum_p M |G| / sum_p M K (FP64 in
  the same block order; the GPU fuses each multiply-add, so a cancelling
  signed sum differs in its last bits), weights within 1e-13 relative,
  covariances the oracle's covariance_matrix (src/utils.f90:745-764) within
  1e-12 of the largest entry.
* Drawn multiplicities: every sample holds n_draw = int(P * supersampling)
  draws; the same seed draws the same, another seed differently; the pooled
  draws are uniform over the blocks (chi-square below its 99.9 % point); the
  returned multiplicities reproduce the outputs through the oracle.
* Seeded random configurations (block, bin and sample counts, empty
  block-bins, given or drawn multiplicities, with or without weights).
* End to end: a correlation with patches, block_sums, and the bootstrap with
  unit multiplicities returns the correlation's own g_t, g_x (1e-12).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
SHEAR = 0


def blocks_of(seed, P, nb):
    r = np.random.default_rng(seed)
    k = r.uniform(1, 5, (P, nb))
    k[0, 3] = 0.0   # an empty block-bin
    gt = r.normal(0, 0.01, (P, nb)) * k
    gt[0, 3] = 0.0
    gx = r.normal(0, 0.01, (P, nb)) * k
    gx[0, 3] = 0.0
    return dict(g_t=gt, g_x=gx, k=k, weight=r.uniform(10, 20, P))


def close(g, o, blocks, M):
    Mf = M.astype(np.float64)
    with np.errstate(invalid="ignore", divide="ignore"):
        for key in ("g_t", "g_x"):
            scale = (Mf @ np.abs(blocks[key])) / (Mf @ blocks["k"])
            scale = np.where(np.isfinite(scale), scale, 0.0)
            assert np.all(np.abs(g[key] - o[key]) <= 1e-12 * scale + 1e-300), key
    np.testing.assert_allclose(g["weights"], o["weights"], rtol=1e-13)


def test_bootstrap_given_multiplicities(tp, orc, ctx):
    P, nb, n_boot = 37, 20, 300
    blocks = blocks_of(1, P, nb)
    r = np.random.default_rng(2)
    M = r.integers(0, 3, (n_boot, P)).astype(np.int32)
    M[5] = 0   # a sample with nothing drawn: K = 0 everywhere
    g = tp.bootstrap(blocks, n_boot, multiplicity=M, ctx=ctx)
    o = orc.bootstrap(blocks, M)
    close(g, o, blocks, M)
    assert np.all(g["g_t"][5] == 0)
    np.testing.assert_array_equal(g["multiplicity"], M)
    for key, src in (("cov_t", "g_t"), ("cov_x", "g_x")):
        c = orc.covariance(o[src])
        assert np.max(np.abs(g[key] - c)) <= 1e-12 * np.max(np.abs(c)), key


def test_bootstrap_drawn_multiplicities(tp, orc, ctx):
    P, nb, n_boot = 64, 10, 2000
    blocks = blocks_of(3, P, nb)
    a = tp.bootstrap(blocks, n_boot, seed=11, supersampling=1.5, ctx=ctx)
    b = tp.bootstrap(blocks, n_boot, seed=11, supersampling=1.5, ctx=ctx)
    c = tp.bootstrap(blocks, n_boot, seed=12, supersampling=1.5, ctx=ctx)
    M = a["multiplicity"]
    assert np.all(M.sum(axis=1) == int(P * 1.5))
    np.testing.assert_array_equal(M, b["multiplicity"])
    assert not np.array_equal(M, c["multiplicity"])
    counts = M.sum(axis=0).astype(float)
    expect = M.sum() / P
    chi2 = np.sum((counts - expect) ** 2 / expect)
    assert chi2 < 120.0      # 99.9 % point of chi-square with 63 degrees of freedom is 103
    close(a, orc.bootstrap(blocks, M), blocks, M)


@pytest.mark.parametrize("seed", range(10))
def test_bootstrap_random_configurations(tp, orc, ctx, seed):
    """Seeded random cases against the oracle at the bars above: 1 to 700
    blocks, 1 to 63 bins, 2 to 500 samples, empty block-bins, given
    multiplicities (with an all-zero sample) or drawn ones (supersampling
    0.5 to 2), with or without block weights."""
    r = np.random.default_rng(7000 + seed)
    P = int(r.choice([1, 2, 37, 128, 700]))
    nb = int(r.choice([1, 5, 30, 63]))
    n_boot = int(r.choice([2, 10, 500]))
    k = r.uniform(1, 5, (P, nb))
    k[r.random((P, nb)) < 0.1] = 0.0
    blocks = dict(g_t=r.normal(0, 0.01, (P, nb)) * k, g_x=r.normal(0, 0.01, (P, nb)) * k, k=k)
    if r.random() < 0.7:
        blocks["weight"] = r.uniform(10, 20, P)
    given = r.random() < 0.5
    print(f"\nseed {seed}: {P} blocks, {nb} bins, {n_boot} samples, given multiplicities {given}", flush=True)
    if given:
        M = r.integers(0, 3, (n_boot, P)).astype(np.int32)
        M[int(r.integers(0, n_boot))] = 0
        g = tp.bootstrap(blocks, n_boot, multiplicity=M, ctx=ctx)
        np.testing.assert_array_equal(g["multiplicity"], M)
    else:
        ss = float(r.uniform(0.5, 2.0))
        if int(P * ss) < 1:
            ss = 1.0   # at least one draw per sample (fewer is refused)
        g = tp.bootstrap(blocks, n_boot, seed=int(r.integers(0, 2**31)), supersampling=ss, ctx=ctx)
        M = g["multiplicity"]
        assert np.all(M.sum(axis=1) == int(P * ss)) and np.all(M >= 0)
    o = orc.bootstrap(blocks, M)
    if "weight" in blocks:
        close(g, o, blocks, M)
    else:
        Mf = M.astype(np.float64)
        with np.errstate(invalid="ignore", divide="ignore"):
            for key in ("g_t", "g_x"):
                scale = np.nan_to_num((Mf @ np.abs(blocks[key])) / (Mf @ blocks["k"]))
                assert np.all(np.abs(g[key] - o[key]) <= 1e-12 * scale + 1e-300), key
    for key, src in (("cov_t", "g_t"), ("cov_x", "g_x")):
        c = orc.covariance(o[src])
        assert np.max(np.abs(g[key] - c)) <= 1e-12 * np.max(np.abs(c)) + 1e-300, key


def test_bootstrap_end_to_end_unit_multiplicities(tp, orc, ctx):
    r = np.random.default_rng(8)
    n_a, n_b, P = 20000, 20000, 16
    f = np.float32
    lens = dict(x=r.uniform(0, 120, n_a).astype(f), y=r.uniform(0, 120, n_a).astype(f), z=np.zeros(n_a, f),
                val=r.normal(0, 0.02, n_a).astype(f), w=r.uniform(0.5, 1.5, n_a).astype(f))
    src = dict(x=r.uniform(0, 120, n_b).astype(f), y=r.uniform(0, 120, n_b).astype(f), z=np.ones(n_b, f),
               e1=r.normal(0, 0.27, n_b).astype(f), e2=r.normal(0, 0.27, n_b).astype(f),
               w=r.uniform(0.5, 1.5, n_b).astype(f), m=r.normal(0, 0.05, n_b).astype(f))
    pa = (np.floor(lens["x"] / 30) + 4 * np.floor(lens["y"] / 30)).astype(np.int32)
    pb = (np.floor(src["x"] / 30) + 4 * np.floor(src["y"] / 30)).astype(np.int32)
    bins = tp.create_bins(1.0, 20.0, 10, tp.LOG)
    res = tp.correlate(SHEAR, lens, src, bins, patch_A=pa, patch_B=pb, n_patch_A=P, n_patch_B=P, real_bytes=8,
                       ctx=ctx)
    blocks = tp.block_sums(res, pa, lens["w"], P)
    g = tp.bootstrap(blocks, 1, multiplicity=np.ones((1, P), np.int32), ctx=ctx)
    np.testing.assert_allclose(g["g_t"][0], res.g_t, rtol=1e-12)
    np.testing.assert_allclose(g["g_x"][0], res.g_x, rtol=1e-12)
    np.testing.assert_allclose(g["weights"][0], lens["w"].astype(np.float64).sum(), rtol=1e-12)
    # and 200 drawn resamples scatter around it
    d = tp.bootstrap(blocks, 200, seed=5, ctx=ctx)
    sd = np.sqrt(np.diag(d["cov_t"]))
    assert np.all(np.abs(d["g_t"].mean(axis=0) - res.g_t) < 4 * sd / np.sqrt(200) + 1e-12)


def test_bootstrap_errors(tp, ctx):
    blocks = blocks_of(4, 5, 4)
    with pytest.raises(tp.TpcfError):
        tp.bootstrap(blocks, 0, ctx=ctx)
    with pytest.raises(tp.TpcfError):
        tp.bootstrap(blocks, 3, supersampling=0.0, ctx=ctx)
