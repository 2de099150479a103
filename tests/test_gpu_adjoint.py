"""The adjoint (tpcf_adjoint_*) and its torch.autograd.Function on the GPU.

Checked against
  * an independent FP64 gradient: the forward written as a brute-force FP64
    torch computation on the CPU over every pair (bins decided by the oracle's
    bin definition), differentiated by torch autograd;
  * central finite differences of the library's own FP64 forward with steps
    of 1e-4;
on catalogues with z masks (the reversed z test of the source-side pass),
zero lens weights, m != 0, and a kappa auto-correlation through autograd
(the autograd forward records the lens-side pass, its backward replays it);
the recorded form against the one-call adjoint; 12 seeded random
configurations (sizes, sides, bins, spacings, kind, precision, masks).
Tolerance: FP32 path within 2e-5 of the largest |gradient| of each array
(per-pair terms are FP32); FP64 path within 1e-10; finite differences within
1e-6 relative (+1e-9 absolute)."""
import importlib

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
SHEAR, KAPPA = 0, 1


def catalogue(seed, n_a, n_b, side, kind, masks=True):
    r = np.random.default_rng(seed)
    f = np.float32
    lens = dict(x=r.uniform(0, side, n_a).astype(f), y=r.uniform(0, side, n_a).astype(f),
                z=(r.uniform(-0.2, 1.0, n_a) if masks else np.zeros(n_a)).astype(f),
                val=r.normal(0, 0.02, n_a).astype(f), w=r.uniform(0.5, 1.5, n_a).astype(f))
    if masks:
        lens["w"][::17] = 0.0
    if kind == SHEAR:
        src = dict(x=r.uniform(0, side, n_b).astype(f), y=r.uniform(0, side, n_b).astype(f),
                   z=(r.uniform(-0.2, 1.2, n_b) if masks else np.ones(n_b)).astype(f),
                   e1=r.normal(0, 0.27, n_b).astype(f), e2=r.normal(0, 0.27, n_b).astype(f),
                   w=r.uniform(0.5, 1.5, n_b).astype(f), m=r.normal(0, 0.05, n_b).astype(f))
    else:
        lens.pop("z")
        src = dict(x=r.uniform(0, side, n_b).astype(f), y=r.uniform(0, side, n_b).astype(f),
                   kappa_E=r.normal(0, 0.02, n_b).astype(f), kappa_B=r.normal(0, 0.02, n_b).astype(f),
                   w=r.uniform(0.5, 1.5, n_b).astype(f))
    return lens, src


def brute_grads(orc, kind, lens, src, edges, gbar1, gbar2):
    """FP64 torch autograd through a brute-force forward over all pairs."""
    n_bin = len(edges) - 1
    P = {k: torch.tensor(np.asarray(v, np.float64), requires_grad=k not in ("x", "y", "z"))
         for k, v in lens.items()}
    S = {k: torch.tensor(np.asarray(v, np.float64), requires_grad=k not in ("x", "y", "z")) for k, v in src.items()}
    xa, ya = np.asarray(lens["x"], np.float64), np.asarray(lens["y"], np.float64)
    xb, yb = np.asarray(src["x"], np.float64), np.asarray(src["y"], np.float64)
    dx = xa[:, None] - xb[None, :]
    dy = ya[:, None] - yb[None, :]
    d2 = np.fma(dx, dx, dy * dy) if hasattr(np, "fma") else dx * dx + dy * dy
    b = orc.bin_of_d2(d2.ravel(), edges).reshape(d2.shape)
    keep = b >= 0
    if kind == SHEAR:
        za, zb = np.asarray(lens["z"]), np.asarray(src["z"])
        keep &= (za[:, None] >= 0) & (zb[None, :] >= 0) & (za[:, None] <= zb[None, :])
    ii, jj = np.nonzero(keep)
    bb = torch.tensor(b[ii, jj], dtype=torch.long)
    ii_t, jj_t = torch.tensor(ii), torch.tensor(jj)
    wa, a = P["w"][ii_t], P["val"][ii_t] * P["w"][ii_t]
    wb = S["w"][jj_t]
    if kind == SHEAR:
        ddx, ddy = torch.tensor(dx[ii, jj]), torch.tensor(dy[ii, jj])
        r2 = ddx * ddx + ddy * ddy
        c, s = (ddx * ddx - ddy * ddy) / r2, 2 * ddx * ddy / r2
        e1, e2 = S["e1"][jj_t], S["e2"][jj_t]
        t1 = -(e1 * c + e2 * s) * wb * a
        t2 = -(-e1 * s + e2 * c) * wb * a
        k = wa * wb * (1 + S["m"][jj_t])
    else:
        t1 = a * wb * S["kappa_E"][jj_t]
        t2 = a * wb * S["kappa_B"][jj_t]
        k = wa * wb
    G = torch.zeros(n_bin, dtype=torch.float64).index_add(0, bb, t1)
    X = torch.zeros(n_bin, dtype=torch.float64).index_add(0, bb, t2)
    K = torch.zeros(n_bin, dtype=torch.float64).index_add(0, bb, k)
    Ks = torch.where(K > 0, K, torch.ones_like(K))
    g1 = torch.where(K > 0, G / Ks, torch.zeros_like(G))
    g2 = torch.where(K > 0, X / Ks, torch.zeros_like(X))
    loss = (torch.tensor(gbar1) * g1).sum() + (torch.tensor(gbar2) * g2).sum()
    loss.backward()
    lens_g = {k: P[k].grad.numpy() for k in ("val", "w")}
    src_keys = ("e1", "e2", "w", "m") if kind == SHEAR else ("kappa_E", "kappa_B", "w")
    src_g = {k: S[k].grad.numpy() for k in src_keys}
    return {"lens": lens_g, "src": src_g}, g1.detach().numpy(), g2.detach().numpy()


def zero_masked(kind, lens, src, g):
    """Objects masked by z < 0 get zero gradient by definition."""
    if kind == SHEAR:
        for k in g["lens"]:
            g["lens"][k] = np.where(np.asarray(lens["z"]) < 0, 0.0, g["lens"][k])
        for k in g["src"]:
            g["src"][k] = np.where(np.asarray(src["z"]) < 0, 0.0, g["src"][k])
    return g


@pytest.mark.parametrize("kind", [SHEAR, KAPPA])
@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_adjoint_matches_fp64_autograd(tp, orc, ctx, kind, precision):
    lens, src = catalogue(300 + kind, 500, 600, 40.0, kind)
    bins = tp.create_bins(0.5, 12.0, 9, tp.LOG)
    r = np.random.default_rng(9)
    gbar1, gbar2 = r.normal(size=bins.n_bin), r.normal(size=bins.n_bin)
    ref, _, _ = brute_grads(orc, kind, lens, src, bins.bin_edges, gbar1, gbar2)
    ref = zero_masked(kind, lens, src, ref)
    dt = np.float32 if precision == "f32" else np.float64
    got = tp.adjoint(kind, {k: v.astype(dt) for k, v in lens.items()}, {k: v.astype(dt) for k, v in src.items()},
                     bins, gbar1, gbar2, precision=precision, ctx=ctx)
    tol = 2e-5 if precision == "f32" else 1e-10
    for side in ("lens", "src"):
        for k, want in ref[side].items():
            have = got[side][k]
            scale = np.abs(want).max()
            err = np.abs(have - want).max()
            print(f"{side}.{k}: max |grad| {scale:.3e}, max err {err:.3e}", flush=True)
            assert err <= tol * scale + 1e-300, (side, k, err, scale)


@pytest.mark.parametrize("seed", range(12))
def test_adjoint_random_configurations(tp, orc, ctx, seed):
    """Seeded random cases against the FP64 autograd gradient at the bars
    above: 1 to 700 lenses (chunks of 1, 7, 64 and 65), 1 to 800 sources,
    field sides of 5' to 60', 1 to 63 bins of linear (x_min = 0 or not),
    logarithmic or square-root spacing, both kinds, FP32 and FP64, with and
    without z masks and zero lens weights.  With a single object on a side
    (or a single one left by the masks) g = G / K does not depend on its
    weight: the weight gradient is the difference of two equal terms and the
    FP64 reference holds only its rounding.  An array whose reference
    gradient lies below 1e-12 of the call's largest |gradient| is therefore
    held to the bar against that largest |gradient|; every other array to the
    bar against its own."""
    r = np.random.default_rng(5000 + seed)
    kind = int(r.integers(0, 2))
    precision = "f32" if r.random() < 0.5 else "f64"
    n_a = int(r.choice([1, 7, 64, 65, 300, 700, 700]))
    n_b = int(r.choice([1, 5, 200, 800, 800, 800]))
    side = float(r.uniform(5.0, 60.0))
    masks = bool(r.random() < 0.6)
    lens, src = catalogue(5100 + seed, n_a, n_b, side, kind, masks=masks)
    spacing = [tp.LIN, tp.LOG, tp.SQRT][int(r.integers(0, 3))]
    x_min = 0.0 if spacing == tp.LIN and r.random() < 0.4 else float(r.uniform(0.2, 2.0))
    x_max = float(r.uniform(x_min + 1.0, 1.2 * side + 2.0))
    n_bin = int(r.integers(1, 64))
    bins = tp.create_bins(x_min, x_max, n_bin, spacing)
    print(f"\nseed {seed}: kind {kind}, {precision}, {n_a} x {n_b} on {side:.1f}', masks {masks}, "
          f"{n_bin} bins {x_min:.2f}-{x_max:.2f} spacing {spacing}", flush=True)
    gbar1, gbar2 = r.normal(size=n_bin), r.normal(size=n_bin)
    ref, _, _ = brute_grads(orc, kind, lens, src, bins.bin_edges, gbar1, gbar2)
    ref = zero_masked(kind, lens, src, ref)
    dt = np.float32 if precision == "f32" else np.float64
    got = tp.adjoint(kind, {k: v.astype(dt) for k, v in lens.items()}, {k: v.astype(dt) for k, v in src.items()},
                     bins, gbar1, gbar2, precision=precision, ctx=ctx)
    tol = 2e-5 if precision == "f32" else 1e-10
    top = max(np.abs(v).max() for sd in ref.values() for v in sd.values())
    for side_ in ("lens", "src"):
        for k, want in ref[side_].items():
            err = np.abs(got[side_][k] - want).max()
            scale = np.abs(want).max()
            if scale <= 1e-12 * top:
                scale = top
            print(f"{side_}.{k}: max |grad| {np.abs(want).max():.3e} (scale {scale:.3e}), max err {err:.3e}")
            assert err <= tol * scale + 1e-300, (side_, k, err, scale)


def test_adjoint_finite_differences(tp, ctx):
    kind = SHEAR
    lens, src = catalogue(311, 300, 400, 30.0, kind)
    lens = {k: v.astype(np.float64) for k, v in lens.items()}
    src = {k: v.astype(np.float64) for k, v in src.items()}
    bins = tp.create_bins(0.5, 10.0, 8, tp.LOG)
    r = np.random.default_rng(3)
    gbar1, gbar2 = r.normal(size=bins.n_bin), r.normal(size=bins.n_bin)

    def loss(L, S):
        g = tp.correlate(kind, L, S, bins, precision="f64", real_bytes=8, ctx=ctx)
        return float(np.dot(gbar1, g.g_t) + np.dot(gbar2, g.g_x))

    got = tp.adjoint(kind, lens, src, bins, gbar1, gbar2, precision="f64", ctx=ctx)
    h = 1e-4
    for side, arrs in (("lens", lens), ("src", src)):
        for k in got[side]:
            for idx in r.choice(len(arrs[k]), size=4, replace=False):
                if side == "lens" and arrs["z"][idx] < 0 or side == "src" and arrs["z"][idx] < 0:
                    continue
                plus = {kk: vv.copy() for kk, vv in arrs.items()}
                minus = {kk: vv.copy() for kk, vv in arrs.items()}
                plus[k][idx] += h
                minus[k][idx] -= h
                if side == "lens":
                    fd = (loss(plus, src) - loss(minus, src)) / (2 * h)
                else:
                    fd = (loss(lens, plus) - loss(lens, minus)) / (2 * h)
                ad = got[side][k][idx]
                print(f"{side}.{k}[{idx}]: adjoint {ad:.9e} fd {fd:.9e}", flush=True)
                assert abs(ad - fd) <= 1e-6 * abs(fd) + 1e-9, (side, k, idx, ad, fd)


def test_autograd_function_chi2(tp, orc, ctx):
    ag = importlib.import_module("2pcf_amd.autograd")
    dev = torch.device("cuda:0")
    lens, src = catalogue(320, 400, 500, 35.0, SHEAR)
    bins = tp.create_bins(0.5, 10.0, 8, tp.LOG)
    L = {k: torch.tensor(v, device=dev, requires_grad=k in ("val", "w")) for k, v in lens.items()}
    S = {k: torch.tensor(v, device=dev, requires_grad=k in ("e1", "e2", "w", "m")) for k, v in src.items()}
    g_t, g_x = ag.shear_scalar(L, S, bins, ctx=ctx)
    model = torch.zeros_like(g_t)
    sigma = torch.full_like(g_t, 1e-3)
    loss = ag.chi2(g_t, model, sigma) + ag.chi2(g_x, model, sigma)
    loss.backward()
    # the same gradient from the brute-force FP64 reference with gbar = dchi2/dg = 2 g / sigma^2
    gb1 = (2 * g_t / sigma ** 2).detach().cpu().numpy()
    gb2 = (2 * g_x / sigma ** 2).detach().cpu().numpy()
    ref, r1, r2 = brute_grads(orc, SHEAR, lens, src, bins.bin_edges, gb1, gb2)
    ref = zero_masked(SHEAR, lens, src, ref)
    np.testing.assert_allclose(g_t.detach().cpu().numpy(), r1, rtol=1e-5, atol=1e-9)
    for side, T in (("lens", L), ("src", S)):
        for k, want in ref[side].items():
            have = T[k].grad.detach().cpu().double().numpy()
            assert np.abs(have - want).max() <= 2e-5 * np.abs(want).max(), (side, k)


def test_autograd_kappa_auto_correlation(tp, orc, ctx):
    # the same catalogue on both sides: autograd adds both contributions
    ag = importlib.import_module("2pcf_amd.autograd")
    dev = torch.device("cuda:0")
    r = np.random.default_rng(12)
    n = 600
    x, y = r.uniform(0, 40, n).astype(np.float32), r.uniform(0, 40, n).astype(np.float32)
    k, w = r.normal(0, 0.02, n).astype(np.float32), r.uniform(0.5, 1.5, n).astype(np.float32)
    kb = r.normal(0, 0.02, n).astype(np.float32)
    bins = tp.create_bins(0.5, 12.0, 9, tp.LOG)
    X, Y = torch.tensor(x, device=dev), torch.tensor(y, device=dev)
    K = torch.tensor(k, device=dev, requires_grad=True)
    W = torch.tensor(w, device=dev, requires_grad=True)
    KB = torch.tensor(kb, device=dev)
    xi_E, xi_B = ag.kappa_kappa(dict(x=X, y=Y, val=K, w=W), dict(x=X, y=Y, kappa_E=K, kappa_B=KB, w=W), bins, ctx=ctx)
    gb = r.normal(size=bins.n_bin)
    (torch.tensor(gb, device=dev) * xi_E).sum().backward()
    # FP64 reference with shared parameters
    Kt = torch.tensor(k.astype(np.float64), requires_grad=True)
    Wt = torch.tensor(w.astype(np.float64), requires_grad=True)
    xd, yd = x.astype(np.float64), y.astype(np.float64)
    dx, dy = xd[:, None] - xd[None, :], yd[:, None] - yd[None, :]
    b = orc.bin_of_d2((dx * dx + dy * dy).ravel(), bins.bin_edges).reshape(dx.shape)
    ii, jj = np.nonzero(b >= 0)
    bb = torch.tensor(b[ii, jj])
    it, jt = torch.tensor(ii), torch.tensor(jj)
    G = torch.zeros(bins.n_bin, dtype=torch.float64).index_add(0, bb, Kt[it] * Wt[it] * Wt[jt] * Kt[jt])
    Wb = torch.zeros(bins.n_bin, dtype=torch.float64).index_add(0, bb, Wt[it] * Wt[jt])
    (torch.tensor(gb) * (G / Wb)).sum().backward()
    for have, want in ((K.grad, Kt.grad), (W.grad, Wt.grad)):
        h, wnt = have.detach().cpu().double().numpy(), want.numpy()
        assert np.abs(h - wnt).max() <= 2e-5 * np.abs(wnt).max()


@pytest.mark.parametrize("kind", [SHEAR, KAPPA])
def test_record_replay_matches_adjoint(tp, ctx, kind):
    """tpcf_adjoint_record_* + tpcf_adjoint_replay: the recorded forward
    equals the plain forward (FP32 path: within 1e-5 of the largest |g|), and
    each replay equals tpcf_adjoint_* for the same upstream gradient (the same
    passes: within 1e-9 of each array's largest |gradient|), also after other
    calls on the same context have reused its buffers, and twice in a row."""
    dev = torch.device("cuda:0")
    lens, src = catalogue(330 + kind, 500, 600, 40.0, kind)
    bins = tp.create_bins(0.5, 12.0, 9, tp.LOG)
    dl = {k: torch.tensor(v, device=dev) for k, v in lens.items()}
    ds = {k: torch.tensor(v, device=dev) for k, v in src.items()}
    torch.cuda.synchronize()
    g1, g2, tape = tp.adjoint_record(kind, {k: v.data_ptr() for k, v in dl.items()},
                                     {k: v.data_ptr() for k, v in ds.items()}, bins, precision="f32", ctx=ctx,
                                     n_A=len(lens["x"]), n_B=len(src["x"]))
    fwd = tp.correlate(kind, lens, src, bins, real_bytes=8, ctx=ctx)
    for have, want in ((g1, fwd.g_t), (g2, fwd.g_x)):
        assert np.abs(have - want).max() <= 1e-5 * np.abs(want).max() + 1e-12
    r = np.random.default_rng(21)
    lk = ("val", "w")
    sk = ("e1", "e2", "w", "m") if kind == SHEAR else ("kappa_E", "kappa_B", "w")
    for trial in range(3):
        gbar1, gbar2 = r.normal(size=bins.n_bin), r.normal(size=bins.n_bin)
        bufs = {"lens": {k: torch.full((len(lens["x"]),), 7.0, dtype=torch.float64, device=dev) for k in lk},
                "src": {k: torch.full((len(src["x"]),), 7.0, dtype=torch.float64, device=dev) for k in sk}}
        torch.cuda.synchronize()
        tp.adjoint_replay(tape, gbar1, gbar2, {s: {k: b.data_ptr() for k, b in d.items()} for s, d in bufs.items()})
        ref = tp.adjoint(kind, lens, src, bins, gbar1, gbar2, precision="f32", ctx=ctx)
        for side in ("lens", "src"):
            for k, want in ref[side].items():
                have = bufs[side][k].cpu().numpy()
                assert np.abs(have - want).max() <= 1e-9 * np.abs(want).max() + 1e-300, (trial, side, k)
    tape.close()


def test_record_needs_device_inputs(tp, ctx):
    import ctypes as C
    N = importlib.import_module("2pcf_amd._native")
    lens, src = catalogue(340, 50, 60, 10.0, KAPPA)
    bins = tp.create_bins(0.5, 5.0, 4, tp.LOG)
    p = lambda a: a.ctypes.data_as(C.c_void_p)   # noqa: E731
    h = C.c_void_p()
    rc = N.lib().tpcf_adjoint_record_kappa_f32(ctx.handle, 50, p(lens["x"]), p(lens["y"]), p(lens["val"]),
                                               p(lens["w"]), 60, p(src["x"]), p(src["y"]), p(src["kappa_E"]),
                                               p(src["kappa_B"]), p(src["w"]), C.byref(bins.c), 0, None, None,
                                               C.byref(h))
    assert rc == N.ERR_ARG and not h.value
    assert b"TPCF_DEVICE_INPUTS" in N.lib().tpcf_last_error(ctx.handle)
