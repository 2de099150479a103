"""Seeded synthetic catalogues for the benchmark workloads and the tests.

Flat square patches in arcmin.  Positions are either uniform or drawn from a
lognormal density field: white noise on an n x n pixel grid smoothed by a
Gaussian of the given FWHM, scaled to unit variance (G), density
proportional to exp(G - 1/2); pixels are drawn with that probability and the
position is jittered uniformly inside the pixel.  Patch labels are k-means
(Lloyd) clusters of the lens positions, the source labels come from the same
centroids (nearest centroid).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


def lognormal_density(rng: np.random.Generator, npix: int, side: float, fwhm: float) -> np.ndarray:
    g = rng.standard_normal((npix, npix)).astype(np.float32)
    pix = side / npix
    sigma_pix = fwhm / (2.0 * np.sqrt(2.0 * np.log(2.0))) / pix
    k = np.fft.fftfreq(npix).astype(np.float32)
    kern = np.exp(-2.0 * (np.pi * sigma_pix) ** 2 * (k[:, None] ** 2 + k[None, :] ** 2)).astype(np.float32)
    gs = np.fft.irfft2(np.fft.rfft2(g) * kern[:, : npix // 2 + 1], s=g.shape).astype(np.float32)
    gs -= gs.mean()
    gs /= gs.std()
    dens = np.exp(gs - 0.5).astype(np.float64)
    return dens / dens.sum()


def sample_positions(rng: np.random.Generator, n: int, side: float, density: np.ndarray | None):
    if density is None:
        return (rng.uniform(0.0, side, n).astype(np.float32), rng.uniform(0.0, side, n).astype(np.float32))
    npix = density.shape[0]
    counts = rng.multinomial(n, density.ravel())
    idx = rng.permutation(np.repeat(np.arange(npix * npix, dtype=np.int64), counts))
    iy, ix = np.divmod(idx, npix)
    pix = side / npix
    x = ((ix + rng.random(n)) * pix).astype(np.float32)
    y = ((iy + rng.random(n)) * pix).astype(np.float32)
    return x, y


def kmeans_centroids(rng: np.random.Generator, x, y, k: int, iters: int = 10, sample: int = 200_000) -> np.ndarray:
    n = len(x)
    sel = rng.choice(n, size=min(n, sample), replace=False)
    pts = np.stack([x[sel], y[sel]], axis=1).astype(np.float64)
    cen = pts[rng.choice(len(pts), size=k, replace=False)].copy()
    for _ in range(iters):
        lab = assign(pts[:, 0], pts[:, 1], cen)
        for c in range(k):
            m = lab == c
            if m.any():
                cen[c] = pts[m].mean(axis=0)
    return cen


def assign(x, y, cen: np.ndarray) -> np.ndarray:
    """Label of the nearest centroid."""
    from scipy.spatial import cKDTree
    _, lab = cKDTree(cen).query(np.stack([np.asarray(x, np.float64), np.asarray(y, np.float64)], axis=1),
                                workers=-1)
    return lab.astype(np.int32)


def ellipticities(rng: np.random.Generator, n: int, sigma: float = 0.27):
    e1 = rng.normal(0.0, sigma, n)
    e2 = rng.normal(0.0, sigma, n)
    bad = e1 * e1 + e2 * e2 >= 1.0
    while bad.any():   # redraw until |e| < 1
        e1[bad] = rng.normal(0.0, sigma, bad.sum())
        e2[bad] = rng.normal(0.0, sigma, bad.sum())
        bad = e1 * e1 + e2 * e2 >= 1.0
    return e1.astype(np.float32), e2.astype(np.float32)


@dataclass
class Workload:
    name: str
    kind: int              # 0 shear, 1 kappa
    lens: dict
    src: dict
    x_min: float
    x_max: float
    n_bin: int
    n_patch: int
    patch_A: np.ndarray | None
    patch_B: np.ndarray | None
    auto: bool = False
    spherical: bool = False   # x, y = longitude, latitude and x_min, x_max in radians


def make_workload(name: str, n_a: int, n_b: int, side: float, kind: int, x_min: float, x_max: float, n_bin: int,
                  n_patch: int, seed: int, clustered: bool = False, npix: int = 2048, fwhm: float = 10.0,
                  auto: bool = False, precision=np.float32, spherical: bool = False) -> Workload:
    rng = np.random.default_rng(seed)
    dens = lognormal_density(rng, npix, side, fwhm) if clustered else None
    xa, ya = sample_positions(rng, n_a, side, dens)
    lens = dict(x=xa, y=ya, val=rng.normal(0.0, 0.02, n_a).astype(np.float32),
                w=rng.uniform(0.5, 1.5, n_a).astype(np.float32), z=np.zeros(n_a, dtype=np.float32))
    if auto:
        # kappa-kappa auto-correlation: the same catalogue on both sides
        kb = rng.normal(0.0, 0.02, n_a).astype(np.float32)
        src = dict(x=xa, y=ya, kappa_E=lens["val"], kappa_B=kb, w=lens["w"])
    else:
        xb, yb = sample_positions(rng, n_b, side, dens)
        wb = rng.uniform(0.5, 1.5, n_b).astype(np.float32)
        if kind == 0:
            e1, e2 = ellipticities(rng, n_b)
            src = dict(x=xb, y=yb, z=np.ones(n_b, dtype=np.float32), e1=e1, e2=e2, w=wb,
                       m=np.zeros(n_b, dtype=np.float32))
        else:
            src = dict(x=xb, y=yb, kappa_E=rng.normal(0.0, 0.02, n_b).astype(np.float32),
                       kappa_B=rng.normal(0.0, 0.02, n_b).astype(np.float32), w=wb)
    pa = pb = None
    if n_patch > 1:
        cen = kmeans_centroids(rng, xa, ya, n_patch)
        pa = assign(xa, ya, cen)
        pb = pa if auto else assign(src["x"], src["y"], cen)
    if spherical:
        # the patch laid on the sphere around (longitude pi, latitude 0): one
        # arcminute of x or y is one arcminute of longitude or latitude
        rad = np.pi / (180.0 * 60.0)
        for d in (lens, src):
            d["x"] = (np.pi + (d["x"].astype(np.float64) - side / 2) * rad).astype(np.float32)
            d["y"] = ((d["y"].astype(np.float64) - side / 2) * rad).astype(np.float32)
        x_min, x_max = x_min * rad, x_max * rad
    if precision != np.float32:
        lens = {k: v.astype(precision) for k, v in lens.items()}
        src = {k: v.astype(precision) for k, v in src.items()}
    return Workload(name, kind, lens, src, x_min, x_max, n_bin, n_patch, pa, pb, auto, spherical)


# The benchmark configurations (sizes, patch sides in arcmin, bins, seeds).
CONFIGS = {
    "kk_100k_auto": dict(n_a=100_000, n_b=100_000, side=600.0, kind=1, x_min=1.0, x_max=100.0, n_bin=20,
                         n_patch=16, seed=1001, auto=True, precision=np.float64),
    "sk_1Mx2M_lognormal_f64": dict(n_a=1_000_000, n_b=2_000_000, side=1800.0, kind=0, x_min=0.5, x_max=250.0,
                                   n_bin=30, n_patch=64, seed=1002, clustered=True, npix=2048,
                                   precision=np.float64),
    "sk_20Mx20M_lognormal_f32": dict(n_a=20_000_000, n_b=20_000_000, side=3000.0, kind=0, x_min=0.5, x_max=250.0,
                                     n_bin=30, n_patch=128, seed=1003, clustered=True, npix=4096),
    "sk_5Mx5M_uniform_f32": dict(n_a=5_000_000, n_b=5_000_000, side=1200.0, kind=0, x_min=1.0, x_max=200.0,
                                 n_bin=20, n_patch=1, seed=1005),
    # the 20M x 20M field on the sphere (50 x 50 degrees around longitude pi, latitude 0)
    "sk_20Mx20M_sphere": dict(n_a=20_000_000, n_b=20_000_000, side=3000.0, kind=0, x_min=0.5, x_max=250.0,
                              n_bin=30, n_patch=128, seed=1003, clustered=True, npix=4096, spherical=True),
}


def mock_fields(n_fields: int = 256, device=None, seed: int = 1004, side: float = 600.0) -> list:
    """The batched form's mock catalogues, generated on the GPU: per field a
    kappa catalogue and a shear catalogue with sizes log-uniform in
    [5e4, 2e6] (numpy PCG64(seed)), uniform on a side x side patch (torch
    generator seeded with seed); kappa ~ N(0, 0.02) (kappa_B independent),
    e1, e2 ~ N(0, 0.27), w ~ U[0.5, 1.5], lens z = 0, source z = 1, and 16
    patches as 4 x 4 squares.  Per field two catalogues: kappa-kappa (the
    kappa catalogue on both sides) and kappa x gamma_t.  Each entry: kind,
    lens / src dicts of torch tensors, patch_A, patch_B (int32 tensors),
    n_patch."""
    import torch
    dev = device or torch.device("cuda:0")
    rng = np.random.default_rng(seed)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    q = side / 4.0

    def uni(n, lo, hi):
        return torch.rand(n, generator=gen, device=dev) * (hi - lo) + lo

    def patch(x, y):
        return (torch.clamp((x / q).int(), 0, 3) + 4 * torch.clamp((y / q).int(), 0, 3)).int()

    cats = []
    for _ in range(n_fields):
        nk, ng = (int(np.exp(rng.uniform(np.log(5e4), np.log(2e6)))) for _ in range(2))
        kx, ky = uni(nk, 0, side), uni(nk, 0, side)
        kap = torch.randn(nk, generator=gen, device=dev) * 0.02
        kw = uni(nk, 0.5, 1.5)
        kz = torch.zeros(nk, device=dev)
        gx, gy = uni(ng, 0, side), uni(ng, 0, side)
        e1 = torch.randn(ng, generator=gen, device=dev) * 0.27
        e2 = torch.randn(ng, generator=gen, device=dev) * 0.27
        gw = uni(ng, 0.5, 1.5)
        gz = torch.ones(ng, device=dev)
        kb = torch.randn(nk, generator=gen, device=dev) * 0.02
        pk, pg = patch(kx, ky), patch(gx, gy)
        lens = dict(x=kx, y=ky, z=kz, val=kap, w=kw)
        cats.append(dict(kind=1, lens=lens, src=dict(x=kx, y=ky, kappa_E=kap, kappa_B=kb, w=kw), patch_A=pk,
                         patch_B=pk, n_patch=16))
        cats.append(dict(kind=0, lens=lens, src=dict(x=gx, y=gy, z=gz, e1=e1, e2=e2, w=gw), patch_A=pk,
                         patch_B=pg, n_patch=16))
    return cats


def batch_desc(cat: dict) -> dict:
    """A catalogue of mock_fields as a batch descriptor (device addresses)."""
    return dict(kind=cat["kind"], lens={k: v.data_ptr() for k, v in cat["lens"].items()},
                src={k: v.data_ptr() for k, v in cat["src"].items()}, n_A=cat["lens"]["x"].numel(),
                n_B=cat["src"]["x"].numel(), patch_A=cat["patch_A"].data_ptr(), patch_B=cat["patch_B"].data_ptr(),
                n_patch_A=cat["n_patch"], n_patch_B=cat["n_patch"])


def config(name: str, scale: float = 1.0) -> Workload:
    """A benchmark workload; ``scale`` < 1 shrinks the catalogue sizes and the
    patch side together (constant density, so per-object neighbour counts keep)."""
    c = dict(CONFIGS[name])
    if scale != 1.0:
        c["n_a"] = max(1, int(c["n_a"] * scale))
        c["n_b"] = max(1, int(c["n_b"] * scale))
        c["side"] = c["side"] * np.sqrt(scale)
        if "npix" in c:
            c["npix"] = max(64, int(c["npix"] * np.sqrt(scale)))
    return make_workload(name, **c)
