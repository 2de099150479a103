"""Precomputed marks through a marks file, with the GPU correlation.

Marks of unit-value, unit-weight lenses (the marks files the reference's
estimator expects, src/main.f90:1069-1076) come from tpcf_shear_f32, go
through write_marks_fits / read_marks_fits (FP32 and FP64 files) and are
re-weighted with the real lens values and weights (tpcf_from_marks,
src/main.f90:1017-1139).  The result equals a direct GPU correlation with
those values within 1e-6 of the per-pair scale sigma_e <|val|> (the FP32 path
rounds each pair's terms to FP32 in both; an FP32 file adds the rounding of
each lens' marks).
"""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
SHEAR = 0


@pytest.mark.parametrize("double", [True, False])
def test_marks_file_estimator_matches_direct(tp, ctx, tmp_path, double):
    mio = importlib.import_module("2pcf_amd.marks_io")
    r = np.random.default_rng(23)
    f = np.float32
    n_a, n_b = 20000, 30000
    lens = dict(x=r.uniform(0, 100, n_a).astype(f), y=r.uniform(0, 100, n_a).astype(f), z=np.zeros(n_a, f),
                val=r.normal(0, 0.02, n_a).astype(f), w=r.uniform(0.5, 1.5, n_a).astype(f))
    lens["w"][::13] = 0
    src = dict(x=r.uniform(0, 100, n_b).astype(f), y=r.uniform(0, 100, n_b).astype(f), z=np.ones(n_b, f),
               e1=r.normal(0, 0.27, n_b).astype(f), e2=r.normal(0, 0.27, n_b).astype(f),
               w=r.uniform(0.5, 1.5, n_b).astype(f), m=r.normal(0, 0.05, n_b).astype(f))
    bins = tp.create_bins(0.5, 15.0, 8, tp.LOG)
    unit = dict(lens, val=np.ones(n_a, f), w=np.ones(n_a, f))
    m = tp.correlate(SHEAR, unit, src, bins, marks=True, real_bytes=8, ctx=ctx)
    field = mio.MarksField(x=lens["x"], y=lens["y"], w=m.marks_weight, xi_E_marks=np.asarray(m.xi_E_marks).T,
                           xi_B_marks=np.asarray(m.xi_B_marks).T, K_marks=np.asarray(m.K_marks).T)
    mf = mio.MarksFile(n_bin=8, spacing=1, theta_min=0.5, theta_max=15.0, bin_centers=bins.bin_centers,
                       bin_edges=bins.bin_edges, fields=[field])
    path = str(tmp_path / "marks.fits")
    mio.write_marks_fits(path, mf, double_precision=double)
    back = mio.read_marks_fits(path)
    L = type("L", (), dict(x=lens["x"], y=lens["y"], val=lens["val"], w=lens["w"]))
    res = mio.tpcf_from_marks(back, [L])
    d = tp.correlate(SHEAR, lens, src, bins, real_bytes=8, ctx=ctx)
    scale = 0.27 * np.mean(np.abs(lens["val"]))
    for key in ("g_t", "g_x"):
        assert np.all(np.abs(res["fields"][0][key] - getattr(d, key)) <= 1e-6 * scale), key
