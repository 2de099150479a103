"""The headline workload at its full size: 20M kappa lenses x 20M shear
sources on a 3000' x 3000' lognormal-clustered patch, 30 log bins 0.5'-250',
128 x 128 patches, FP32 inputs.  Checked by what holds at any size:

  * a seeded sample of 1024 lenses against all 20M sources: pair counts per
    bin and per patch pair bit-exact, sums within 1e-6 (the FP32 bar), against
    the FP64 oracle (cell-pruned) on the same inputs.  The source cells,
    segments and their classes are those of the full catalogue;
  * the full 20M x 20M call with marks: the per-lens marks (which carry the
    lens factors, src/tpcf.f90:354-358) summed over the lenses equal the
    per-bin sums (W_b = sum_i w_marks_ib, K_b, G_t and G_x likewise, and
    sum_i marks_weight_i = sum_b W_b), which reach the output by another route
    (wave sums into the per-bin table, per-patch-pair atomics, reduction):
    within 1e-9 of the sum of |terms|; every bin is populated and the pair
    count per lens is the one the sample predicts within its sampling error.
"""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
SHEAR = 0


@pytest.fixture(scope="module")
def wl():
    synth = importlib.import_module("2pcf_amd.synth")
    w = synth.config("sk_20Mx20M_lognormal_f32")
    print(f"\nworkload {w.name}: {len(w.lens['x'])} lenses x {len(w.src['x'])} sources", flush=True)
    return w


def test_fullsize_lens_sample_against_oracle(tp, orc, ctx, wl):
    r = np.random.default_rng(2024)
    sel = np.sort(r.choice(len(wl.lens["x"]), 1024, replace=False))
    lens = {k: np.ascontiguousarray(v[sel]) for k, v in wl.lens.items()}
    pa = np.ascontiguousarray(wl.patch_A[sel])
    bins = tp.create_bins(wl.x_min, wl.x_max, wl.n_bin, tp.LOG)
    g = tp.correlate(SHEAR, lens, wl.src, bins, patch_A=pa, patch_B=wl.patch_B, n_patch_A=wl.n_patch,
                     n_patch_B=wl.n_patch, real_bytes=8, ctx=ctx)
    print("GPU pairs:", int(g.n_pairs.sum()), "pair kernel ms:", g.stats["ms_pairs"], flush=True)
    o = orc.correlate(SHEAR, lens, wl.src, bins.bin_edges, pa, wl.patch_B, wl.n_patch, wl.n_patch, patches=True)
    np.testing.assert_array_equal(g.n_pairs, o["n_pairs"])
    np.testing.assert_array_equal(g.pp_n, o["pp_n"])
    assert np.all(o["n_pairs"] > 0)
    for gk, ok in (("sum_theta", "theta"), ("sum_w", "w"), ("sum_k", "k")):
        assert np.max(np.abs(getattr(g, gk) - o[ok]) / np.abs(o[ok])) <= 1e-6, gk
    for gk, ok, ak in (("sum_gt", "gt", "gt_abs"), ("sum_gx", "gx", "gx_abs")):
        assert np.max(np.abs(getattr(g, gk) - o[ok]) / o[ak]) <= 1e-6, gk
    for gk, ok in (("pp_w", "pp_w"), ("pp_k", "pp_k"), ("pp_theta", "pp_theta")):
        den = np.where(o[ok] != 0, np.abs(o[ok]), 1.0)
        assert np.max(np.abs(getattr(g, gk) - o[ok]) / den) <= 1e-6, gk
    wl.sample_pairs_per_lens = float(o["n_pairs"].sum()) / len(sel)


def test_fullsize_marks_contract_to_bin_sums(tp, ctx, wl):
    import torch
    dev = torch.device("cuda:0")
    dl = {k: torch.from_numpy(np.ascontiguousarray(v)).to(dev) for k, v in wl.lens.items()}
    ds = {k: torch.from_numpy(np.ascontiguousarray(v)).to(dev) for k, v in wl.src.items()}
    pa = torch.from_numpy(wl.patch_A).to(dev)
    pb = torch.from_numpy(wl.patch_B).to(dev)
    torch.cuda.synchronize()
    bins = tp.create_bins(wl.x_min, wl.x_max, wl.n_bin, tp.LOG)
    n = len(wl.lens["x"])
    g = tp.correlate(SHEAR, {k: v.data_ptr() for k, v in dl.items()}, {k: v.data_ptr() for k, v in ds.items()},
                     bins, patch_A=pa.data_ptr(), patch_B=pb.data_ptr(), n_patch_A=wl.n_patch,
                     n_patch_B=wl.n_patch, real_bytes=8, device_inputs=True, ctx=ctx, n_A=n, n_B=n, marks=True,
                     marks_device=True)
    print("pairs:", int(g.n_pairs.sum()), "pair kernel ms:", g.stats["ms_pairs"], flush=True)
    assert np.all(g.n_pairs > 0)
    # marks are (n_bin, n_A) views of [n_A][n_bin] arrays
    W = g.w_marks.sum(dim=1).cpu().numpy()
    K = g.K_marks.sum(dim=1).cpu().numpy()
    Gt = g.xi_E_marks.sum(dim=1).cpu().numpy()
    Gx = g.xi_B_marks.sum(dim=1).cpu().numpy()
    Gt_abs = g.xi_E_marks.abs().sum(dim=1).cpu().numpy()
    Gx_abs = g.xi_B_marks.abs().sum(dim=1).cpu().numpy()
    np.testing.assert_allclose(W, g.sum_w, rtol=1e-9)
    np.testing.assert_allclose(K, g.sum_k, rtol=1e-9)
    assert np.all(np.abs(Gt - g.sum_gt) <= 1e-9 * Gt_abs)
    assert np.all(np.abs(Gx - g.sum_gx) <= 1e-9 * Gx_abs)
    np.testing.assert_allclose(float(g.marks_weight.sum()), float(g.sum_w.sum()), rtol=1e-9)
    # the marks of 1024 sampled lenses of this full call (its cells, chunks
    # and heavy-chunk work items) against the oracle's marks of those lenses
    # (src/tpcf.f90:354-358): W, K within 1e-6 relative, the signed marks
    # within 1e-6 of each lens-bin's sum of |terms|
    import oracle as orc
    r = np.random.default_rng(2025)
    rows = np.sort(r.choice(n, 1024, replace=False))
    ls = {k: np.ascontiguousarray(v[rows]) for k, v in wl.lens.items()}
    o = orc.correlate(SHEAR, ls, wl.src, bins.bin_edges, marks=True)
    bound = orc.correlate(1, dict(x=ls["x"], y=ls["y"], val=np.abs(ls["val"]), w=ls["w"]),
                          dict(x=wl.src["x"], y=wl.src["y"], kappa_E=np.abs(wl.src["e1"]) + np.abs(wl.src["e2"]),
                               kappa_B=np.zeros(n), w=wl.src["w"]), bins.bin_edges, marks=True)["xi_e"]
    assert tuple(g.w_marks.shape) == (wl.n_bin, n) and rows.max() < n   # the gather below stays in bounds
    rt = torch.from_numpy(rows).to(dev)
    for gk, ok in (("K_marks", "k_marks"), ("w_marks", "w_marks")):
        a = getattr(g, gk)[:, rt].T.cpu().numpy()
        den = np.where(o[ok] != 0, np.abs(o[ok]), 1.0)
        err = np.max(np.abs(a - o[ok]) / den)
        print(gk, "sampled lenses, max relative error:", err, flush=True)
        assert err <= 1e-6, gk
    for gk, ok in (("xi_E_marks", "xi_e"), ("xi_B_marks", "xi_b")):
        a = getattr(g, gk)[:, rt].T.cpu().numpy()
        err = np.max(np.abs(a - o[ok]) / np.maximum(bound, 1e-300))
        print(gk, "sampled lenses, max error / sum |terms|:", err, flush=True)
        assert err <= 1e-6, gk
    np.testing.assert_allclose(g.marks_weight[rt].cpu().numpy(), o["marks_weight"], rtol=1e-6)
    ppl = getattr(wl, "sample_pairs_per_lens", None)
    if ppl is not None:
        # the 1024-lens sample predicts the full pair count to a few per cent
        per_lens = float(g.n_pairs.sum()) / n
        assert abs(per_lens - ppl) <= 0.1 * ppl, (per_lens, ppl)


def test_fullsize_adjoint_record_matches_plain_forward(tp):
    """The differentiable form at its full size (5M kappa points x 5M shear
    sources uniform on 1200' x 1200', 20 log bins 1'-200'): the forward the
    recorded adjoint contracts from its per-lens sums equals the plain
    correlation's per-bin outputs (the same FP32 per-lens partial sums, added
    in another order: within 1e-9 of each bin's sum of |terms| / K, bounded by
    a kappa correlation of |val| against |s1| + |s2|), for kappa-kappa and
    shear-scalar."""
    import torch
    synth = importlib.import_module("2pcf_amd.synth")
    wl = synth.config("sk_5Mx5M_uniform_f32")
    dev = torch.device("cuda:0")
    n = len(wl.lens["x"])
    bins = tp.create_bins(wl.x_min, wl.x_max, wl.n_bin, tp.LOG)
    ctx = tp.Context(0)
    kap = dict(x=wl.lens["x"], y=wl.lens["y"], val=wl.lens["val"], w=wl.lens["w"])
    cases = ((1, kap, dict(x=wl.lens["x"], y=wl.lens["y"], kappa_E=wl.lens["val"],
                           kappa_B=np.ascontiguousarray(wl.src["e2"] * 0.1), w=wl.lens["w"])),
             (SHEAR, dict(kap, z=wl.lens["z"]), wl.src))
    for kind, lens, src in cases:
        dl = {k: torch.from_numpy(np.ascontiguousarray(v)).to(dev) for k, v in lens.items()}
        ds = {k: torch.from_numpy(np.ascontiguousarray(v)).to(dev) for k, v in src.items()}
        torch.cuda.synchronize()
        pl = {k: v.data_ptr() for k, v in dl.items()}
        ps = {k: v.data_ptr() for k, v in ds.items()}
        g = tp.correlate(kind, pl, ps, bins, real_bytes=8, device_inputs=True, ctx=ctx, n_A=n, n_B=n)
        g1, g2, tape = tp.adjoint_record(kind, pl, ps, bins, ctx=ctx, n_A=n, n_B=n)
        tape.close()
        s1, s2 = ("e1", "e2") if kind == SHEAR else ("kappa_E", "kappa_B")
        va = dl["val"].abs()
        sa = ds[s1].abs() + ds[s2].abs()
        zb = torch.zeros_like(sa)
        torch.cuda.synchronize()
        bound = tp.correlate(1, dict(x=pl["x"], y=pl["y"], val=va.data_ptr(), w=pl["w"]),
                             dict(x=ps["x"], y=ps["y"], kappa_E=sa.data_ptr(), kappa_B=zb.data_ptr(), w=ps["w"]),
                             bins, real_bytes=8, device_inputs=True, ctx=ctx, n_A=n, n_B=n)
        scale = bound.sum_gt / g.sum_k
        print("kind", kind, "pairs", int(g.n_pairs.sum()), flush=True)
        assert np.all(g.n_pairs > 0)
        for have, want in ((g1, g.g_t), (g2, g.g_x)):
            err = np.abs(have - want) / scale
            print("max error / (sum |terms| / K):", float(err.max()), flush=True)
            assert np.all(err <= 1e-9), kind
    ctx.close()
