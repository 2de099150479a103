"""Input/output files of the Fortran host program tests/fortran/fortran_driver.f90."""
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "2pcf_amd", "fortran", "fortran_driver")


def write_input(path, kind, spacing, n_bin, x_min, x_max, p, lens, src, patch_A=None, patch_B=None, n_pa=1, n_pb=1):
    f4 = np.float32
    n_a, n_b = len(lens["x"]), len(src["x"])
    with open(path, "wb") as f:
        f.write(np.array([kind, spacing, n_bin, n_a, n_b, n_pa, n_pb], np.int32).tobytes())
        f.write(np.array([x_min, x_max, p], f4).tobytes())
        for k in ("x", "y", "z", "val", "w"):
            f.write(np.asarray(lens.get(k, np.zeros(n_a)), f4).tobytes())
        keys = ("x", "y", "z", "e1", "e2", "w", "m") if kind == 0 else ("x", "y", "kappa_E", "kappa_B", "w")
        for k in keys:
            f.write(np.asarray(src.get(k, np.zeros(n_b)), f4).tobytes())
        if n_pa * n_pb > 1:
            f.write(np.asarray(patch_A, np.int32).tobytes())
            f.write(np.asarray(patch_B, np.int32).tobytes())


def read_output(path, n_bin, n_a, n_pa=1, n_pb=1):
    raw = open(path, "rb").read()
    pos = 0

    def take(dtype, count):
        nonlocal pos
        a = np.frombuffer(raw, dtype, count, pos)
        pos += a.nbytes
        return a

    out = {}
    for k in ("theta", "mean_theta", "g_t", "g_x", "weights"):
        out[k] = take(np.float32, n_bin)
    out["n_pairs"] = take(np.int64, n_bin)
    if n_a > 0:
        for k in ("xi_E_marks", "xi_B_marks", "K_marks", "w_marks"):
            out[k] = take(np.float32, n_bin * n_a).reshape(n_a, n_bin)   # Fortran (n_bin, n_A)
    out["marks_weight"] = take(np.float32, n_a)
    if n_pa * n_pb > 1:
        m = n_bin * n_pa * n_pb
        out["pp_n"] = take(np.int64, m).reshape(n_pb, n_pa, n_bin)
        for k in ("pp_theta", "pp_gt", "pp_gx", "pp_w", "pp_k"):
            out[k] = take(np.float64, m).reshape(n_pb, n_pa, n_bin)
    assert pos == len(raw), (pos, len(raw))
    return out
