"""The adjoint on the sphere (TPCF_SPHERICAL): one-call, recorded and autograd.

The source-side pass runs with the catalogues' roles swapped while the shear
angle stays the one at the lens (the streamed object of that pass); these
tests check it against an independent FP64 gradient: the forward written as a
brute-force torch computation over every pair on the CPU, with the
separation and cos 2phi, sin 2phi of each pair from the oracle's restatement
of distance_sphere / angles_sphere (src/utils.f90:512-528, :618-639), bins
decided on squared chords against the oracle's chord edges, differentiated by
torch autograd.  Fields at high latitude (where the angle at the lens and the
angle at the source differ most) and across the lon = 0 meridian, with z
masks, zero lens weights and m != 0.
Tolerance: 1e-9 of each array's largest |gradient| (the geometry and every
sum are FP64 for both input precisions)."""
import importlib

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
SHEAR, KAPPA = 0, 1


def sph_catalogue(seed, n_a, n_b, kind, lon0, lat0, half, dtype=np.float64):
    r = np.random.default_rng(seed)

    def pos(n):
        lon = lon0 + r.uniform(-half, half, n) / np.cos(lat0)
        lat = lat0 + r.uniform(-half, half, n)
        return np.mod(lon, 2 * np.pi), lat

    xa, ya = pos(n_a)
    xb, yb = pos(n_b)
    lens = dict(x=xa, y=ya, z=r.uniform(-0.2, 1.0, n_a), val=r.normal(0, 0.02, n_a), w=r.uniform(0.5, 1.5, n_a))
    lens["w"][::17] = 0.0
    if kind == SHEAR:
        src = dict(x=xb, y=yb, z=r.uniform(-0.2, 1.2, n_b), e1=r.normal(0, 0.27, n_b), e2=r.normal(0, 0.27, n_b),
                   w=r.uniform(0.5, 1.5, n_b), m=r.normal(0, 0.05, n_b))
    else:
        lens.pop("z")
        src = dict(x=xb, y=yb, kappa_E=r.normal(0, 0.02, n_b), kappa_B=r.normal(0, 0.02, n_b),
                   w=r.uniform(0.5, 1.5, n_b))
    return ({k: v.astype(dtype) for k, v in lens.items()}, {k: v.astype(dtype) for k, v in src.items()})


def unit(lon, lat):
    lon, lat = np.asarray(lon, np.float64), np.asarray(lat, np.float64)
    return np.stack([np.cos(lat) * np.cos(lon), np.cos(lat) * np.sin(lon), np.sin(lat)], axis=-1)


def brute_sph_grads(orc, kind, lens, src, edges, gbar1, gbar2):
    n_bin = len(edges) - 1
    P = {k: torch.tensor(np.asarray(v, np.float64), requires_grad=k not in ("x", "y", "z")) for k, v in lens.items()}
    S = {k: torch.tensor(np.asarray(v, np.float64), requires_grad=k not in ("x", "y", "z")) for k, v in src.items()}
    ua, ub = unit(lens["x"], lens["y"]), unit(src["x"], src["y"])
    diff = ua[:, None, :] - ub[None, :, :]
    c2 = np.einsum("ijk,ijk->ij", diff, diff)
    e2c = orc.chord_edges_squared(edges)
    b = np.searchsorted(e2c, c2, side="right") - 1
    b = np.where(c2 == e2c[-1], n_bin - 1, b)
    keep = (c2 >= e2c[0]) & (c2 <= e2c[-1])
    if kind == SHEAR:
        za, zb = np.asarray(lens["z"]), np.asarray(src["z"])
        keep &= (za[:, None] >= 0) & (zb[None, :] >= 0) & (za[:, None] <= zb[None, :])
    ii, jj = np.nonzero(keep)
    bb = torch.tensor(b[ii, jj], dtype=torch.long)
    it, jt = torch.tensor(ii), torch.tensor(jj)
    wa, a = P["w"][it], P["val"][it] * P["w"][it]
    wb = S["w"][jt]
    if kind == SHEAR:
        _, cc, ss = orc.sph_pair(lens["x"][ii], lens["y"][ii], src["x"][jj], src["y"][jj])
        c, s = torch.tensor(cc), torch.tensor(ss)
        e1, e2 = S["e1"][jt], S["e2"][jt]
        t1 = -(e1 * c + e2 * s) * wb * a
        t2 = -(-e1 * s + e2 * c) * wb * a
        k = wa * wb * (1 + S["m"][jt])
    else:
        t1 = a * wb * S["kappa_E"][jt]
        t2 = a * wb * S["kappa_B"][jt]
        k = wa * wb
    G = torch.zeros(n_bin, dtype=torch.float64).index_add(0, bb, t1)
    X = torch.zeros(n_bin, dtype=torch.float64).index_add(0, bb, t2)
    K = torch.zeros(n_bin, dtype=torch.float64).index_add(0, bb, k)
    Ks = torch.where(K > 0, K, torch.ones_like(K))
    g1 = torch.where(K > 0, G / Ks, torch.zeros_like(G))
    g2 = torch.where(K > 0, X / Ks, torch.zeros_like(X))
    ((torch.tensor(gbar1) * g1).sum() + (torch.tensor(gbar2) * g2).sum()).backward()
    src_keys = ("e1", "e2", "w", "m") if kind == SHEAR else ("kappa_E", "kappa_B", "w")
    g = {"lens": {k: P[k].grad.numpy() for k in ("val", "w")}, "src": {k: S[k].grad.numpy() for k in src_keys}}
    if kind == SHEAR:   # objects masked by z < 0 get zero gradient by definition
        for kk in g["lens"]:
            g["lens"][kk] = np.where(np.asarray(lens["z"]) < 0, 0.0, g["lens"][kk])
        for kk in g["src"]:
            g["src"][kk] = np.where(np.asarray(src["z"]) < 0, 0.0, g["src"][kk])
    return g, g1.detach().numpy(), g2.detach().numpy(), len(ii)


def assert_grads(got, ref, tol, label):
    for side in ("lens", "src"):
        for k, want in ref[side].items():
            have = got[side][k]
            scale = np.abs(want).max()
            err = np.abs(have - want).max()
            print(f"{label} {side}.{k}: max |grad| {scale:.3e}, max err {err:.3e}", flush=True)
            assert scale > 0 and err <= tol * scale, (label, side, k, err, scale)


FIELDS = {"high_lat": (2.0, 1.2, 0.04), "meridian": (0.0, -0.5, 0.03)}


@pytest.mark.parametrize("field", sorted(FIELDS))
@pytest.mark.parametrize("kind", [SHEAR, KAPPA])
@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_sphere_adjoint_matches_fp64_autograd(tp, orc, ctx, field, kind, precision):
    lon0, lat0, half = FIELDS[field]
    dt = np.float32 if precision == "f32" else np.float64
    lens, src = sph_catalogue(500 + kind, 500, 600, kind, lon0, lat0, half, dtype=dt)
    bins = tp.create_bins(0.002, 0.04, 8, tp.LOG)
    r = np.random.default_rng(19)
    gbar1, gbar2 = r.normal(size=bins.n_bin), r.normal(size=bins.n_bin)
    ref, _, _, n_pairs = brute_sph_grads(orc, kind, lens, src, bins.bin_edges, gbar1, gbar2)
    print(f"{field} kind {kind} {precision}: {n_pairs} pairs", flush=True)
    got = tp.adjoint(kind, lens, src, bins, gbar1, gbar2, precision=precision, spherical=True, ctx=ctx)
    assert_grads(got, ref, 1e-9, field)


@pytest.mark.parametrize("seed", range(8))
def test_sphere_adjoint_random_configurations(tp, orc, ctx, seed):
    """Seeded random fields against the FP64 gradient at the bar above:
    centres anywhere in longitude and at latitudes up to 1.4 rad (the fixed
    cases above cross lon = 0), half-sides of 0.005 to 0.05 rad, 1 to 400 lenses, 1 to 600
    sources, 1 to 40 logarithmic or linear bins, both kinds and precisions.
    With a single object on a side its weight gradient vanishes identically,
    so an array whose reference gradient lies below 1e-12 of the call's
    largest is held to the bar against that largest."""
    r = np.random.default_rng(9000 + seed)
    kind = int(r.integers(0, 2))
    precision = "f32" if r.random() < 0.5 else "f64"
    dt = np.float32 if precision == "f32" else np.float64
    lon0, lat0, half = float(r.uniform(0, 2 * np.pi)), float(r.uniform(-1.4, 1.4)), float(r.uniform(0.005, 0.05))
    n_a, n_b = int(r.choice([1, 64, 65, 400])), int(r.choice([1, 200, 600, 600]))
    lens, src = sph_catalogue(9100 + seed, n_a, n_b, kind, lon0, lat0, half, dtype=dt)
    spacing = tp.LOG if r.random() < 0.6 else tp.LIN
    x_min = float(r.uniform(0.0005, 0.005)) if spacing == tp.LOG else 0.0
    bins = tp.create_bins(x_min, float(r.uniform(2 * x_min + 0.005, 3 * half)), int(r.integers(1, 41)), spacing)
    print(f"\nseed {seed}: kind {kind}, {precision}, {n_a} x {n_b}, centre ({lon0:.3f}, {lat0:.3f}), half {half:.4f}, "
          f"{bins.n_bin} bins to {bins.x_max:.4f}", flush=True)
    gbar1, gbar2 = r.normal(size=bins.n_bin), r.normal(size=bins.n_bin)
    ref, _, _, n_pairs = brute_sph_grads(orc, kind, lens, src, bins.bin_edges, gbar1, gbar2)
    got = tp.adjoint(kind, lens, src, bins, gbar1, gbar2, precision=precision, spherical=True, ctx=ctx)
    top = max(np.abs(v).max() for sd in ref.values() for v in sd.values())
    for side in ("lens", "src"):
        for k, want in ref[side].items():
            scale = np.abs(want).max()
            if scale <= 1e-12 * top:
                scale = top
            err = np.abs(got[side][k] - want).max()
            print(f"{n_pairs} pairs {side}.{k}: scale {scale:.3e}, max err {err:.3e}")
            assert err <= 1e-9 * scale + 1e-300, (side, k, err, scale)


@pytest.mark.parametrize("kind", [SHEAR, KAPPA])
def test_sphere_autograd_record_replay(tp, orc, ctx, kind):
    """The autograd form on the sphere (record, then replay in backward):
    forward outputs against the brute-force FP64 forward, gradients against
    the one-call adjoint for the same upstream gradient."""
    ag = importlib.import_module("2pcf_amd.autograd")
    dev = torch.device("cuda:0")
    lens, src = sph_catalogue(520 + kind, 400, 500, kind, 2.0, 1.2, 0.04, dtype=np.float64)
    bins = tp.create_bins(0.002, 0.04, 8, tp.LOG)
    diff_l = ("val", "w")
    diff_s = ("e1", "e2", "w", "m") if kind == SHEAR else ("kappa_E", "kappa_B", "w")
    L = {k: torch.tensor(v, device=dev, requires_grad=k in diff_l) for k, v in lens.items()}
    S = {k: torch.tensor(v, device=dev, requires_grad=k in diff_s) for k, v in src.items()}
    fn = ag.shear_scalar if kind == SHEAR else ag.kappa_kappa
    g1, g2 = fn(L, S, bins, ctx=ctx, spherical=True)
    r = np.random.default_rng(23)
    gb1, gb2 = r.normal(size=bins.n_bin), r.normal(size=bins.n_bin)
    ((torch.tensor(gb1, device=dev) * g1).sum() + (torch.tensor(gb2, device=dev) * g2).sum()).backward()
    ref, r1, r2, _ = brute_sph_grads(orc, kind, lens, src, bins.bin_edges, gb1, gb2)
    for have, want in ((g1, r1), (g2, r2)):
        h = have.detach().cpu().numpy()
        assert np.abs(h - want).max() <= 1e-10 * np.abs(want).max()
    got = {"lens": {k: L[k].grad.detach().cpu().numpy() for k in diff_l},
           "src": {k: S[k].grad.detach().cpu().numpy() for k in diff_s}}
    assert_grads(got, ref, 1e-9, "autograd")
    one = tp.adjoint(kind, lens, src, bins, gb1, gb2, precision="f64", spherical=True, ctx=ctx)
    assert_grads(got, one, 1e-12, "replay vs one call")
