"""Argument checks of the entry points that take a binning or re-read values.

* Every entry point that bins pairs refuses hand-built bins whose edges are
  not finite, non-negative and increasing (TPCF_ERR_BINS): the pair kernel's
  bin search assumes increasing squared edges, so such bins would silently
  give wrong counts.  The correlation calls had this check; the batch, the
  adjoint and the recorded adjoint must have it too.
* A batch fixes its masks when it is planned (lenses with w == 0, and for
  shear z < 0 on either side, are dropped then, as the reference skips them,
  src/tpcf.f90:285, :306-309).  A run whose new values would mask a kept
  object must fail (TPCF_ERR_MASK) instead of silently counting it; values
  that keep every mask run as before.
"""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
SHEAR, KAPPA = 0, 1
DEV = torch.device("cuda:0")


def bad_bins(tp):
    good = tp.create_bins(1.0, 30.0, 8, tp.LOG)
    b = type(good.c)()
    C.memmove(C.byref(b), C.byref(good.c), C.sizeof(b))
    b.edges[4] = b.edges[3]          # not increasing
    good.c = b
    return good


def catalogue(seed, n_a, n_b, side=50.0):
    r = np.random.default_rng(seed)
    f = np.float32
    lens = dict(x=r.uniform(0, side, n_a).astype(f), y=r.uniform(0, side, n_a).astype(f),
                z=np.zeros(n_a, f), val=r.normal(0, 0.02, n_a).astype(f), w=r.uniform(0.5, 1.5, n_a).astype(f))
    src = dict(x=r.uniform(0, side, n_b).astype(f), y=r.uniform(0, side, n_b).astype(f), z=np.ones(n_b, f),
               e1=r.normal(0, 0.27, n_b).astype(f), e2=r.normal(0, 0.27, n_b).astype(f),
               w=r.uniform(0.5, 1.5, n_b).astype(f), m=np.zeros(n_b, f))
    dl = {k: torch.from_numpy(v).to(DEV) for k, v in lens.items()}
    ds = {k: torch.from_numpy(v).to(DEV) for k, v in src.items()}
    torch.cuda.synchronize()
    return lens, src, dl, ds


def ptrs(d):
    return {k: v.data_ptr() for k, v in d.items()}


def test_bad_bins_refused_everywhere(tp, ctx):
    bins = bad_bins(tp)
    lens, src, dl, ds = catalogue(1, 500, 600)
    n_a, n_b = len(lens["x"]), len(src["x"])
    with pytest.raises(tp.TpcfError) as e:
        tp.correlate(SHEAR, lens, src, bins, ctx=ctx)
    assert e.value.code == tp._native.ERR_BINS
    with pytest.raises(tp.TpcfError) as e:
        tp.adjoint(SHEAR, lens, src, bins, np.ones(8), np.ones(8), ctx=ctx)
    assert e.value.code == tp._native.ERR_BINS
    with pytest.raises(tp.TpcfError) as e:
        tp.adjoint_record(SHEAR, ptrs(dl), ptrs(ds), bins, ctx=ctx, n_A=n_a, n_B=n_b)
    assert e.value.code == tp._native.ERR_BINS
    with pytest.raises(tp.TpcfError) as e:
        tp.Batch([dict(kind=SHEAR, lens=ptrs(dl), src=ptrs(ds), n_A=n_a, n_B=n_b)], bins, ctx=ctx)
    assert e.value.code == tp._native.ERR_BINS


@pytest.mark.parametrize("graph", [True, False])
def test_batch_mask_breach(tp, ctx, graph):
    bins = tp.create_bins(1.0, 30.0, 8, tp.LOG)
    lens, src, dl, ds = catalogue(2, 800, 900)
    n_a, n_b = len(lens["x"]), len(src["x"])
    b = tp.Batch([dict(kind=SHEAR, lens=ptrs(dl), src=ptrs(ds), n_A=n_a, n_B=n_b)], bins, graph=graph, ctx=ctx)
    first = b.run()[0]
    assert first["n_pairs"].sum() > 0
    # new values that keep every mask: runs
    dl["val"].mul_(2.0)
    torch.cuda.synchronize()
    again = b.run()[0]
    np.testing.assert_array_equal(again["n_pairs"], first["n_pairs"])
    # a kept lens now behind every source would be masked by z_A > z_B only
    # (a per-pair test the kernel applies each run), not a planning mask:
    # still runs; a kept lens with z < 0 or w == 0, or a kept source with
    # z < 0, is a breach
    for side, key, value in (("lens", "z", -1.0), ("lens", "w", 0.0), ("src", "z", -0.5)):
        d = dl if side == "lens" else ds
        old = float(d[key][7])
        d[key][7] = value
        torch.cuda.synchronize()
        with pytest.raises(tp.TpcfError) as e:
            b.run()
        assert e.value.code == tp._native.ERR_MASK, (side, key)
        d[key][7] = old
        torch.cuda.synchronize()
    # restored values run again and give the first result
    dl["val"].mul_(0.5)
    torch.cuda.synchronize()
    last = b.run()[0]
    np.testing.assert_array_equal(last["n_pairs"], first["n_pairs"])
    np.testing.assert_allclose(last["sum_gt"], first["sum_gt"], rtol=1e-12, atol=1e-300)
    b.close()
