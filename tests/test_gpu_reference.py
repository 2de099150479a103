"""The HIP path against the outputs of the reference's own
tangential_shear_tpcf / scalar_kappa_tpcf (src/tpcf.f90:190-545) on the inputs
stored with them in tests/golden/ref_corr.npz: uniform and clustered
catalogues, z masks, zero weights, m != 0, lin / log / ihs / sqrt spacing, a
kappa auto-correlation with x_min = 0 (self pairs), and zero separations in a
shear correlation with x_min = 0 (NaN in the reference's first bin).

Bar (tests/refcases.py): n_pairs identical away from the FP32/FP64 edge band,
weights / mean_theta within FP32 accumulation error, g_t / g_x within 1e-6 of
sum|terms|/W, per-lens marks within 1e-4 relative, NaN where the reference has
NaN."""
import numpy as np
import pytest

import refcases

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("c", range(refcases.n_cases()))
def test_product_matches_reference_outputs(tp, ctx, c):
    kind, sp, lo, hi, nb, p, lens, src, ref = refcases.load(c)
    print(f"\nreference case {c}: kind {kind} spacing {sp} [{lo}, {hi}] n_bin {nb} "
          f"n_A {len(lens['x'])} n_B {len(src['x'])}", flush=True)
    bins = tp.create_bins(lo, hi, nb, sp, p)
    g = tp.correlate(kind, lens, src, bins, precision="f32", marks=True, real_bytes=4, ctx=ctx)
    got = dict(n_pairs=g.n_pairs, weights=g.weights, g_t=g.g_t, g_x=g.g_x, mean_theta=g.mean_theta,
               xi_e=g.xi_E_marks.T, xi_b=g.xi_B_marks.T, k_marks=g.K_marks.T, w_marks=g.w_marks.T,
               marks_weight=g.marks_weight)
    print("pairs per bin:", g.n_pairs.tolist(), flush=True)
    refcases.compare(kind, lens, src, bins.bin_edges, ref, got)
