"""Loader and comparison for the correlation golden vectors written by the
reference's own tangential_shear_tpcf / scalar_kappa_tpcf (src/tpcf.f90:190-545,
tests/golden/ref_corr.npz, kd-tree path, FP32, marks on).

The reference bins in FP32 (distance = sqrt(dx**2 + dy**2), then log10 / asinh
/ sqrt of it, src/utils.f90:101-137) and accumulates in FP32; the product and
the oracle bin on FP64 squared separations and accumulate in FP64.  The two
definitions can only disagree for a pair whose separation lies within a few
FP32 ulps of a nonzero bin edge (such a pair can change bin, or be dropped by the
reference at x_max, src/utils.f90:132-136).  ``compare`` therefore demands:
  * n_pairs equal wherever no pair lies within BAND (relative) of an edge of
    that bin, and |difference| <= the number of such pairs elsewhere;
  * weights, K-normalised g_t / g_x, mean_theta and the per-lens marks within
    FP32 accumulation error of the reference (tolerances below), plus what the
    edge-band pairs can move.
"""
import os

import numpy as np

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
# Relative half-width of the edge band.  The reference's FP32 find_bin_index
# disagrees with the FP64 squared-edge test only within 1e-6 of an edge (the
# widest seen, log spacing up to x = 1024, is 9.5e-7; tests/golden/ref_bins.npz
# holds dense FP32 values around every edge), and its FP32 distance
# sqrt(dx**2 + dy**2) adds at most a few 1e-7.
BAND = 3e-6
RTOL_SUM = 2e-4      # FP32 accumulation in the reference over <= 1e5 terms
RTOL_G = 1e-6        # of sum|terms| for the signed sums G_t, G_x (seen: <= 3e-8)


def n_cases() -> int:
    return int(np.load(os.path.join(GOLD, "ref_corr.npz"))["n_cases"])


def load(c: int):
    """(kind, spacing, x_min, x_max, n_bin, p, lens, src, ref) of case ``c``."""
    g = np.load(os.path.join(GOLD, "ref_corr.npz"))
    kind, sp, lo, hi, nb, p = g[f"case{c}_params"]
    lens, src, ref = {}, {}, {}
    pre = f"case{c}_"
    for key in g.files:
        if not key.startswith(pre):
            continue
        rest = key[len(pre):]
        if rest.startswith("lens_"):
            lens[rest[5:]] = g[key]
        elif rest.startswith("src_"):
            src[rest[4:]] = g[key]
        elif rest.startswith("ref_"):
            ref[rest[4:]] = g[key]
    return int(kind), int(sp), float(lo), float(hi), int(nb), float(p), lens, src, ref


def scan_pairs(kind, lens, src, edges):
    """Brute force in FP64 over the masked pairs.  Returns ``band`` (n_A, n_bin):
    pairs whose separation is within BAND of an edge of that bin (counted for
    both bins of the edge); and ``absum`` (n_bin): the sum over the bin's pairs
    of |s| w_B |val_A| w_A with |s| = sqrt(s1^2 + s2^2), a bound of the sum of
    |terms| of both G components, i.e. the scale of their rounding error."""
    e = np.asarray(edges, np.float64)
    n_bin = len(e) - 1
    e2 = e * e
    xa, ya = lens["x"].astype(np.float64), lens["y"].astype(np.float64)
    xb, yb = src["x"].astype(np.float64), src["y"].astype(np.float64)
    s1, s2 = ("e1", "e2") if kind == 0 else ("kappa_E", "kappa_B")
    smag = np.hypot(src[s1].astype(np.float64), src[s2].astype(np.float64)) * src["w"].astype(np.float64)
    n_a = len(xa)
    cnt = np.zeros((n_a, n_bin), np.int64)
    absum = np.zeros(n_bin)
    for i in range(n_a):
        if lens["w"][i] == 0:
            continue
        d2 = (xa[i] - xb) ** 2 + (ya[i] - yb) ** 2
        d = np.sqrt(d2)
        keep = np.ones(len(xb), bool)
        if kind == 0:
            keep = (lens["z"][i] >= 0) & (src["z"] >= 0) & (lens["z"][i] <= src["z"])
        k = np.searchsorted(e2, d2, side="right") - 1
        k[d2 == e2[-1]] = n_bin - 1
        inb = keep & (k >= 0) & (k < n_bin)
        absum += np.bincount(k[inb], weights=smag[inb], minlength=n_bin) * abs(float(lens["val"][i])) * \
            float(lens["w"][i])
        for j, ej in enumerate(e):
            if ej == 0.0:     # d = 0 = x_min is in the first bin in both definitions
                continue
            m = int((keep & (np.abs(d - ej) <= BAND * ej)).sum())
            if m:
                if j > 0:
                    cnt[i, j - 1] += m
                if j < n_bin:
                    cnt[i, j] += m
    return cnt, absum


def compare(kind, lens, src, edges, ref, got, scan=None):
    """``got`` holds n_pairs, weights, g_t, g_x, mean_theta and optionally the
    marks (n_A, n_bin) of the product or the oracle; ``ref`` the reference's.
    ``scan`` is scan_pairs' result (computed when not given)."""
    band, absum = scan if scan is not None else scan_pairs(kind, lens, src, edges)
    bb = band.sum(axis=0)
    n_ref, n_got = ref["n_pairs"].astype(np.int64), np.asarray(got["n_pairs"], np.int64)
    assert np.all(np.abs(n_ref - n_got) <= bb), (n_ref, n_got, bb)
    # one pair's weight is at most max(w_A) * max(w_B)
    wmax = float(np.max(lens["w"]) * np.max(src["w"]))
    W_ref, W_got = ref["weights"].astype(np.float64), np.asarray(got["weights"], np.float64)
    assert np.all(np.abs(W_ref - W_got) <= RTOL_SUM * np.abs(W_got) + bb * wmax + 1e-30), (W_ref, W_got)
    ok = (n_got > 0) & (bb == 0)
    for key in ("g_t", "g_x"):
        a, b = ref[key].astype(np.float64), np.asarray(got[key], np.float64)
        nan_a, nan_b = np.isnan(a), np.isnan(b)
        assert np.array_equal(nan_a, nan_b), (key, a, b)
        fin = ok & ~nan_a
        # g = G/K: the reference's FP32 error of G is a small multiple of
        # eps32 * sum|terms|, and K ~ W
        with np.errstate(divide="ignore", invalid="ignore"):
            tol = RTOL_G * absum / np.abs(W_got)
        assert np.all(np.abs(a - b)[fin] <= tol[fin]), (key, a[fin], b[fin], tol[fin])
    mt_a, mt_b = ref["mean_theta"].astype(np.float64), np.asarray(got["mean_theta"], np.float64)
    assert np.array_equal(np.isnan(mt_a), np.isnan(mt_b))
    fin = ok & ~np.isnan(mt_a)
    np.testing.assert_allclose(mt_a[fin], mt_b[fin], rtol=RTOL_SUM)
    if "xi_e" in got:
        clean = band == 0
        for key in ("xi_e", "xi_b", "k_marks", "w_marks"):
            a, b = ref[key].astype(np.float64), np.asarray(got[key], np.float64)
            assert a.shape == b.shape, (key, a.shape, b.shape)
            nan_a = np.isnan(a)
            assert np.array_equal(nan_a, np.isnan(b)), key
            m = clean & ~nan_a
            scale = np.abs(b[m]).max() if m.any() else 0.0
            assert np.all(np.abs(a - b)[m] <= 1e-4 * np.abs(b)[m] + 1e-6 * scale + 1e-9), key
        mw_ok = band.sum(axis=1) == 0
        np.testing.assert_allclose(ref["marks_weight"][mw_ok], np.asarray(got["marks_weight"])[mw_ok], rtol=1e-4,
                                   atol=1e-6)
