"""The Fortran host program (tests/fortran/fortran_driver.f90, module
tpcf_gpu) on the GPU: the reference's call sequence create_bins ->
shear_2pcf / kappa_2pcf (and the per-patch-pair extension) returns the same
results as the same call made through ctypes, which the other GPU tests pin to
the FP64 oracle and to the reference's own outputs.  Pair counts and patch-pair
counts must be identical; FP32 outputs and FP64 sums equal to the rounding of
FP64 atomics in no fixed order (1e-6 relative to the bin's scale)."""
import subprocess

import numpy as np
import pytest

import fortran_io as F

pytestmark = pytest.mark.gpu


def uniform(seed, n_a, n_b, side, kind):
    r = np.random.default_rng(seed)
    f = np.float32
    lens = dict(x=r.uniform(0, side, n_a).astype(f), y=r.uniform(0, side, n_a).astype(f),
                z=r.uniform(-0.1, 1.0, n_a).astype(f), val=r.normal(0, .02, n_a).astype(f),
                w=r.uniform(.5, 1.5, n_a).astype(f))
    if kind == 0:
        src = dict(x=r.uniform(0, side, n_b).astype(f), y=r.uniform(0, side, n_b).astype(f),
                   z=r.uniform(-0.1, 1.2, n_b).astype(f), e1=r.normal(0, .27, n_b).astype(f),
                   e2=r.normal(0, .27, n_b).astype(f), w=r.uniform(.5, 1.5, n_b).astype(f),
                   m=r.normal(0, .02, n_b).astype(f))
    else:
        src = dict(x=r.uniform(0, side, n_b).astype(f), y=r.uniform(0, side, n_b).astype(f),
                   kappa_E=r.normal(0, .02, n_b).astype(f), kappa_B=r.normal(0, .02, n_b).astype(f),
                   w=r.uniform(.5, 1.5, n_b).astype(f))
    return lens, src


def close(a, b, rel=1e-6):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    assert np.array_equal(np.isnan(a), np.isnan(b))
    m = ~np.isnan(a)
    scale = np.abs(b[m]).max() if m.any() else 0.0
    assert np.all(np.abs(a[m] - b[m]) <= rel * scale + 1e-30)


@pytest.mark.parametrize("kind,patches", [(0, True), (1, False), (0, False), (1, True)])
def test_fortran_host_matches_ctypes(tp, ctx, tmp_path, kind, patches):
    n_a, n_b = 3000, 4000
    lens, src = uniform(70 + kind, n_a, n_b, 90.0, kind)
    n_pa, n_pb = (5, 4) if patches else (1, 1)
    r = np.random.default_rng(5)
    pa = r.integers(0, n_pa, n_a).astype(np.int32) if patches else None
    pb = r.integers(0, n_pb, n_b).astype(np.int32) if patches else None
    x_min, x_max, n_bin, spacing = 0.5, 20.0, 12, tp.LOG
    print(f"\nfortran driver: kind {kind} patches {patches}", flush=True)
    F.write_input(tmp_path / "in", kind, spacing, n_bin, x_min, x_max, 1.0, lens, src, pa, pb, n_pa, n_pb)
    p = subprocess.run([F.DRIVER, str(tmp_path / "in"), str(tmp_path / "out")], capture_output=True, text=True,
                       timeout=120)
    print(p.stdout, p.stderr, flush=True)
    assert p.returncode == 0
    fo = F.read_output(tmp_path / "out", n_bin, n_a, n_pa, n_pb)
    bins = tp.create_bins(x_min, x_max, n_bin, spacing)
    g = tp.correlate(kind, lens, src, bins, marks=True, patch_A=pa, patch_B=pb, n_patch_A=n_pa, n_patch_B=n_pb,
                     real_bytes=4, ctx=ctx)
    np.testing.assert_array_equal(fo["n_pairs"], g.n_pairs)
    np.testing.assert_array_equal(fo["theta"], bins.bin_centers)
    for k in ("mean_theta", "g_t", "g_x", "weights"):
        close(fo[k], getattr(g, k))
    for k in ("xi_E_marks", "xi_B_marks", "K_marks", "w_marks"):
        close(fo[k], getattr(g, k).T)
    close(fo["marks_weight"], g.marks_weight)
    if patches:
        np.testing.assert_array_equal(fo["pp_n"], g.pp_n)
        for k in ("pp_theta", "pp_gt", "pp_gx", "pp_w", "pp_k"):
            close(fo[k], getattr(g, k))


def test_fortran_patch_labels_and_bootstrap(tp, ctx, tmp_path):
    """tests/fortran/fortran_patches.f90: partition_data_gpu and
    block_bootstrap_gpu from Fortran give the ctypes results (1-based labels,
    the reference's (n_boot, n_bin) bootstrap layout)."""
    import os
    r = np.random.default_rng(41)
    n, k, n_blk, n_bin, n_boot, seed, ss = 50000, 24, 24, 10, 400, 77, 1.0
    cen = r.uniform(0, 100, (8, 2))
    lab0 = r.integers(0, 8, n)
    x = (cen[lab0, 0] + r.normal(0, 6, n)).astype(np.float32)
    y = (cen[lab0, 1] + r.normal(0, 6, n)).astype(np.float32)
    w = r.uniform(0.5, 1.5, n).astype(np.float32)
    bk = r.uniform(1, 3, (n_blk, n_bin)).astype(np.float32)
    bt = (r.normal(0, 0.01, (n_blk, n_bin)) * bk).astype(np.float32)
    bx = (r.normal(0, 0.01, (n_blk, n_bin)) * bk).astype(np.float32)
    bw = r.uniform(1, 2, n_blk).astype(np.float32)
    with open(tmp_path / "in", "wb") as f:
        f.write(np.array([n, k, n_blk, n_bin, n_boot], np.int32).tobytes())
        f.write(np.array([seed], np.int64).tobytes())
        f.write(np.array([ss], np.float32).tobytes())
        f.write(x.tobytes() + y.tobytes() + w.tobytes())
        # Fortran (n_bin, n_blk) column-major == C [n_blk][n_bin]
        f.write(bt.tobytes() + bx.tobytes() + bk.tobytes() + bw.tobytes())
    exe = os.path.join(os.path.dirname(F.DRIVER), "fortran_patches")
    p = subprocess.run([exe, str(tmp_path / "in"), str(tmp_path / "out")], capture_output=True, text=True,
                       timeout=120)
    print(p.stdout, p.stderr, flush=True)
    assert p.returncode == 0
    raw = open(tmp_path / "out", "rb").read()
    pos = 0

    def take(dt, count, shape=None, order="C"):
        nonlocal pos
        a = np.frombuffer(raw, dt, count, pos)
        pos += a.nbytes
        return a.reshape(shape, order=order) if shape else a

    lab = take(np.int32, n)
    fcen = take(np.float32, 2 * k, (k, 2), "F")
    gt = take(np.float32, n_boot * n_bin, (n_boot, n_bin), "F")
    gx = take(np.float32, n_boot * n_bin, (n_boot, n_bin), "F")
    ww = take(np.float32, n_boot)
    ct = take(np.float32, n_bin * n_bin, (n_bin, n_bin), "F")
    km = tp.kmeans(x, y, w, k, ctx=ctx)
    np.testing.assert_array_equal(lab, km["labels"] + 1)
    np.testing.assert_allclose(fcen, km["centroids"], rtol=1e-6, atol=1e-5)
    b = tp.bootstrap(dict(g_t=bt, g_x=bx, k=bk, weight=bw), n_boot, seed=seed, supersampling=ss, ctx=ctx)
    close(gt, b["g_t"])
    close(gx, b["g_x"])
    close(ww, b["weights"])
    close(ct, b["cov_t"])
