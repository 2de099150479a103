"""tpcf_shear_f32_noalloc, the allocation-free variant of tpcf_shear_f32.

Called through the C-ABI (ctypes), it must give what tpcf_shear_f32 gives on
the same inputs: pair counts per bin and per patch pair identical, FP64 sums
and marks equal up to the order of FP64 atomics (1e-12 of each output's
scale), on two calls in a row where the second is smaller (its buffers sit
where the first call's were, so anything not cleared would show).  Both host
and device inputs are checked.  The free device memory does not change over
calls of growing size, a call too large for the scratch region is refused
with TPCF_ERR_ALLOC and a message naming the limit, and the context's own
buffers keep working for the allocating entry point afterwards.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
SHEAR = 0
DEV = torch.device("cuda:0")


def catalogue(seed, n_a, n_b, side, n_patch):
    r = np.random.default_rng(seed)
    f = np.float32
    lens = dict(x=r.uniform(0, side, n_a).astype(f), y=r.uniform(0, side, n_a).astype(f),
                z=r.uniform(-0.1, 1.0, n_a).astype(f), val=r.normal(0, 0.02, n_a).astype(f),
                w=r.uniform(0.5, 1.5, n_a).astype(f))
    lens["w"][::97] = 0.0                       # skipped lenses
    src = dict(x=r.uniform(0, side, n_b).astype(f), y=r.uniform(0, side, n_b).astype(f),
               z=r.uniform(-0.1, 1.2, n_b).astype(f), e1=r.normal(0, 0.27, n_b).astype(f),
               e2=r.normal(0, 0.27, n_b).astype(f), w=r.uniform(0.5, 1.5, n_b).astype(f),
               m=r.normal(0, 0.02, n_b).astype(f))
    pa = r.integers(0, n_patch, n_a).astype(np.int32)
    pb = r.integers(0, n_patch, n_b).astype(np.int32)
    return lens, src, pa, pb


def call(tp, ctx, cat, n_patch, bins, device, allocation_free):
    lens, src, pa, pb = cat
    kw = dict(patch_A=pa, patch_B=pb, n_patch_A=n_patch, n_patch_B=n_patch, real_bytes=8, ctx=ctx, marks=True,
              allocation_free=allocation_free)
    if not device:
        return tp.correlate(SHEAR, lens, src, bins, **kw)
    dl = {k: torch.from_numpy(v).to(DEV) for k, v in lens.items()}
    ds = {k: torch.from_numpy(v).to(DEV) for k, v in src.items()}
    kw["patch_A"] = torch.from_numpy(pa).to(DEV)
    kw["patch_B"] = torch.from_numpy(pb).to(DEV)
    torch.cuda.synchronize()
    keep = (dl, ds, kw["patch_A"], kw["patch_B"])
    kw["patch_A"], kw["patch_B"] = kw["patch_A"].data_ptr(), kw["patch_B"].data_ptr()
    r = tp.correlate(SHEAR, {k: v.data_ptr() for k, v in dl.items()}, {k: v.data_ptr() for k, v in ds.items()}, bins,
                     device_inputs=True, n_A=len(lens["x"]), n_B=len(src["x"]), **kw)
    del keep
    return r


def same(a, b):
    np.testing.assert_array_equal(a.n_pairs, b.n_pairs)
    np.testing.assert_array_equal(a.pp_n, b.pp_n)
    assert a.n_pairs.sum() > 0
    for k in ("sum_theta", "sum_gt", "sum_gx", "sum_w", "sum_k", "pp_theta", "pp_gt", "pp_gx", "pp_w", "pp_k"):
        x, y = getattr(a, k), getattr(b, k)
        assert np.all(np.abs(x - y) <= 1e-12 * np.abs(y).max() + 1e-300), k
    for k in ("xi_E_marks", "xi_B_marks", "K_marks", "w_marks", "marks_weight"):
        x, y = np.asarray(getattr(a, k)), np.asarray(getattr(b, k))
        assert np.all(np.abs(x - y) <= 1e-12 * np.abs(y).max() + 1e-300), k


@pytest.mark.parametrize("device", [False, True])
def test_noalloc_matches_allocating_call(tp, device):
    ctx = tp.Context(0)
    bins = tp.create_bins(0.5, 20.0, 14, tp.LOG)
    big = catalogue(31, 60000, 80000, 300.0, 8)
    small = catalogue(32, 7000, 9000, 100.0, 4)
    for cat, n_patch in ((big, 8), (small, 4)):
        print(f"\nnoalloc call: {len(cat[0]['x'])} x {len(cat[1]['x'])}, device inputs {device}", flush=True)
        g = call(tp, ctx, cat, n_patch, bins, device, True)
        ref = call(tp, ctx, cat, n_patch, bins, device, False)
        same(g, ref)
    ctx.close()


def test_noalloc_allocates_nothing_and_refuses_oversize(tp):
    ctx = tp.Context(0)
    bins = tp.create_bins(1.0, 30.0, 10, tp.LOG)
    cats = [catalogue(40 + i, n, n, 200.0, 4) for i, n in enumerate((2000, 20000, 200000))]
    call(tp, ctx, cats[0], 4, bins, False, True)       # first call on the device: symbol lookup
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info(0)[0]
    for cat in cats:
        call(tp, ctx, cat, 4, bins, False, True)
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info(0)[0] == free0
    # beyond the scratch region: refused with a clear error, nothing written
    n = 4_000_000
    r = np.random.default_rng(1)
    f = np.float32
    lens = dict(x=r.uniform(0, 900, n).astype(f), y=r.uniform(0, 900, n).astype(f), z=np.zeros(n, f),
                val=np.zeros(n, f), w=np.ones(n, f))
    src = dict(x=lens["x"], y=lens["y"], z=np.ones(n, f), e1=np.zeros(n, f), e2=np.zeros(n, f), w=np.ones(n, f),
               m=np.zeros(n, f))
    with pytest.raises(tp.TpcfError) as e:
        tp.correlate(SHEAR, lens, src, bins, ctx=ctx, allocation_free=True)
    assert e.value.code == tp._native.ERR_ALLOC and "MiB" in str(e.value)
    # a call that passes the up-front bound but outgrows the region while it
    # runs (its marks, 4 x 150000 lenses x 60 bins in FP64, are 288 MB)
    mid = catalogue(50, 150000, 150000, 200.0, 1)
    with pytest.raises(tp.TpcfError) as e:
        tp.correlate(SHEAR, mid[0], mid[1], tp.create_bins(1.0, 30.0, 60, tp.LOG), ctx=ctx, marks=True,
                     real_bytes=8, allocation_free=True)
    assert e.value.code == tp._native.ERR_ALLOC and "MiB" in str(e.value)
    # the same context still runs both entry points
    a = call(tp, ctx, cats[1], 4, bins, False, False)
    b = call(tp, ctx, cats[1], 4, bins, False, True)
    same(b, a)
    ctx.close()
