"""ctypes front-end of the CPU oracle (``tpcf_oracle.c``) and of the host-core
timing port (``ref_port.c``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` -- always as the checker of the HIP
path, never as the thing measured or shipped.  The product package does not
import this module.

Reference behaviour restated (tilmantroester/2pcf):
  * ``edges``      -- create_bins, src/utils.f90:139-230 (FP64 closed form)
  * ``bin_of_d2``  -- find_bin_index, src/utils.f90:101-137 (squared form)
  * ``correlate``  -- tangential_shear_tpcf src/tpcf.f90:190-378 and
                      scalar_kappa_tpcf src/tpcf.f90:380-545, every pair
                      visited, FP64 geometry, long-double sums
  * ``finalize``   -- src/tpcf.f90:369-376 / :536-543
  * ``ref_port``   -- the reference algorithm (kd-tree, FP32, OpenMP) for timing
  * ``rect_block_labels`` / ``survey_block_sums`` -- the rectangular block
                      scheme of survey_marked_point_block_bootstrap,
                      src/tpcf.f90:1017-1069 (labels through rp_find_bin)
  * ``reference``  -- the reference's own tangential_shear_tpcf /
                      scalar_kappa_tpcf, compiled by ``build_ref.sh`` into
                      ``_ref/libref2pcf.so`` (present only where the reference
                      tree and flang were available when it was built)
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

SHEAR, KAPPA = 0, 1
LIN, LOG, IHS, SQRT = 0, 1, 2, 3

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)


class _OrcArgs(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("n_bin", C.c_int32), ("n_a", C.c_int64), ("n_b", C.c_int64),
        ("xa", _dp), ("ya", _dp), ("za", _dp), ("vala", _dp), ("wa", _dp),
        ("xb", _dp), ("yb", _dp), ("zb", _dp), ("s1b", _dp), ("s2b", _dp), ("wb", _dp), ("mb", _dp),
        ("edges", _dp), ("patch_a", _i32p), ("patch_b", _i32p), ("n_pa", C.c_int32), ("n_pb", C.c_int32),
        ("brute", C.c_int32), ("n_threads", C.c_int32), ("spherical", C.c_int32),
    ]


class _OrcOut(C.Structure):
    _fields_ = [
        ("n_pairs", _i64p), ("theta", _dp), ("gt", _dp), ("gx", _dp), ("w", _dp), ("k", _dp),
        ("xi_e", _dp), ("xi_b", _dp), ("k_marks", _dp), ("w_marks", _dp), ("marks_weight", _dp),
        ("pp_n", _i64p), ("pp_theta", _dp), ("pp_gt", _dp), ("pp_gx", _dp), ("pp_w", _dp), ("pp_k", _dp),
        ("gt_abs", _dp), ("gx_abs", _dp),
    ]


class _RpBins(C.Structure):
    _fields_ = [("x_min", C.c_float), ("x_max", C.c_float), ("bin_size", C.c_float),
                ("ihs_theta", C.c_float), ("n_bin", C.c_int32), ("spacing", C.c_int32)]


class _RpArgs(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("n_a", C.c_int32), ("n_b", C.c_int32),
        ("xa", _fp), ("ya", _fp), ("za", _fp), ("vala", _fp), ("wa", _fp),
        ("xb", _fp), ("yb", _fp), ("zb", _fp), ("s1b", _fp), ("s2b", _fp), ("wb", _fp), ("mb", _fp),
        ("bins", _RpBins), ("n_threads", C.c_int32), ("use_tree", C.c_int32),
    ]


class _RpOut(C.Structure):
    _fields_ = [("mean_theta", _fp), ("g_t", _fp), ("g_x", _fp), ("weights", _fp), ("n_pairs", _i64p),
                ("xi_e", _fp), ("xi_b", _fp), ("k_marks", _fp), ("w_marks", _fp), ("marks_weight", _fp),
                ("tree_seconds", C.c_double)]


REF_LIB_PATH = os.path.join(_HERE, "_ref", "libref2pcf.so")
_ref_lib = None


def build() -> str:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _lib.orc_edges.argtypes = [C.c_int, C.c_float, C.c_float, C.c_int, C.c_float, _dp]
        _lib.orc_edges.restype = C.c_int
        _lib.orc_edges_squared.argtypes = [_dp, C.c_int, _dp]
        _lib.orc_bin_of_d2_array.argtypes = [C.c_int64, _dp, _dp, C.c_int, _i32p]
        _lib.orc_correlate.argtypes = [C.POINTER(_OrcArgs), C.POINTER(_OrcOut)]
        _lib.orc_correlate.restype = C.c_int
        _lib.orc_finalize.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        _lib.orc_chord_edges_squared.argtypes = [_dp, C.c_int, _dp]
        _lib.orc_kmeans.argtypes = [C.c_int64, _dp, _dp, _dp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int,
                                    C.c_int, C.c_int, C.c_int, C.c_double, _i32p, _dp, _dp, C.POINTER(C.c_int64),
                                    C.POINTER(C.c_int64)]
        _lib.orc_kmeans.restype = C.c_int
        _lib.orc_qsort_arg.argtypes = [C.c_int64, _dp, C.POINTER(C.c_int64)]
        _lib.orc_qsort_arg.restype = C.c_int
        _lib.orc_sph_pair.argtypes = [C.c_int64, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        _lib.rp_create_bins.argtypes = [C.c_float, C.c_float, C.c_int, C.c_int, C.c_float, C.POINTER(_RpBins)]
        _lib.rp_create_bins.restype = C.c_int
        _lib.rp_find_bin_array.argtypes = [C.c_int64, _fp, C.POINTER(_RpBins), _i32p]
        _lib.rp_correlate.argtypes = [C.POINTER(_RpArgs), C.POINTER(_RpOut)]
        _lib.rp_correlate.restype = C.c_int
    return _lib


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _ptr(a, t):
    return None if a is None else a.ctypes.data_as(t)


def edges(spacing: int, x_min: float, x_max: float, n_bin: int, ihs_theta: float = 1.0) -> np.ndarray:
    e = np.zeros(n_bin + 1, dtype=np.float64)
    rc = lib().orc_edges(spacing, x_min, x_max, n_bin, ihs_theta, _ptr(e, _dp))
    if rc:
        raise ValueError("ill-defined binning")
    return e


def bin_of_d2(d2, edges_: np.ndarray) -> np.ndarray:
    d2 = _d(np.atleast_1d(d2))
    e = _d(edges_)
    n = len(e) - 1
    e2 = np.zeros_like(e)
    lib().orc_edges_squared(_ptr(e, _dp), n, _ptr(e2, _dp))
    out = np.zeros(len(d2), dtype=np.int32)
    lib().orc_bin_of_d2_array(len(d2), _ptr(d2, _dp), _ptr(e2, _dp), n, _ptr(out, _i32p))
    return out


def correlate(kind: int, lens: dict, src: dict, edges_: np.ndarray, patch_a=None, patch_b=None,
              n_pa: int = 1, n_pb: int = 1, brute: bool = False, marks: bool = False,
              patches: bool = False, n_threads: int | None = None, spherical: bool = False) -> dict:
    """Raw FP64 sums.  ``lens`` has x, y, val, w (+ z for shear); ``src`` has
    x, y, w and e1, e2, m, z (shear) or kappa_E, kappa_B (kappa).  With
    ``spherical`` x, y are longitude and latitude and the edges angles, all in
    radians."""
    L = lib()
    e = _d(edges_)
    n_bin = len(e) - 1
    keep = []

    def arr(dct, key, n):
        v = dct.get(key)
        a = _d(v) if v is not None else np.zeros(n, dtype=np.float64)
        keep.append(a)
        return a

    n_a = len(lens["x"])
    n_b = len(src["x"])
    a = _OrcArgs()
    a.kind = kind
    a.n_bin = n_bin
    a.n_a = n_a
    a.n_b = n_b
    a.xa, a.ya = _ptr(arr(lens, "x", n_a), _dp), _ptr(arr(lens, "y", n_a), _dp)
    a.za = _ptr(arr(lens, "z", n_a), _dp)
    a.vala, a.wa = _ptr(arr(lens, "val", n_a), _dp), _ptr(arr(lens, "w", n_a), _dp)
    a.xb, a.yb = _ptr(arr(src, "x", n_b), _dp), _ptr(arr(src, "y", n_b), _dp)
    a.wb = _ptr(arr(src, "w", n_b), _dp)
    if kind == SHEAR:
        a.zb = _ptr(arr(src, "z", n_b), _dp)
        a.s1b, a.s2b = _ptr(arr(src, "e1", n_b), _dp), _ptr(arr(src, "e2", n_b), _dp)
        a.mb = _ptr(arr(src, "m", n_b), _dp)
    else:
        a.zb = _ptr(arr(src, "z", n_b), _dp)
        a.s1b, a.s2b = _ptr(arr(src, "kappa_E", n_b), _dp), _ptr(arr(src, "kappa_B", n_b), _dp)
        a.mb = _ptr(arr(src, "m", n_b), _dp)
    a.edges = _ptr(e, _dp)
    pa = pb = None
    if patches:
        pa = np.ascontiguousarray(patch_a if patch_a is not None else np.zeros(n_a), dtype=np.int32)
        pb = np.ascontiguousarray(patch_b if patch_b is not None else np.zeros(n_b), dtype=np.int32)
        a.patch_a, a.patch_b = _ptr(pa, _i32p), _ptr(pb, _i32p)
    a.n_pa, a.n_pb = n_pa, n_pb
    a.brute = 1 if brute else 0
    a.n_threads = n_threads or min(16, os.cpu_count() or 1)
    a.spherical = 1 if spherical else 0

    res = {
        "n_pairs": np.zeros(n_bin, dtype=np.int64),
        "theta": np.zeros(n_bin), "gt": np.zeros(n_bin), "gx": np.zeros(n_bin),
        "w": np.zeros(n_bin), "k": np.zeros(n_bin), "gt_abs": np.zeros(n_bin), "gx_abs": np.zeros(n_bin),
    }
    o = _OrcOut()
    o.n_pairs = _ptr(res["n_pairs"], _i64p)
    for key in ("theta", "gt", "gx", "w", "k", "gt_abs", "gx_abs"):
        setattr(o, key, _ptr(res[key], _dp))
    if marks:
        for key in ("xi_e", "xi_b", "k_marks", "w_marks"):
            res[key] = np.zeros((n_a, n_bin))
            setattr(o, key, _ptr(res[key], _dp))
        res["marks_weight"] = np.zeros(n_a)
        o.marks_weight = _ptr(res["marks_weight"], _dp)
    if patches:
        res["pp_n"] = np.zeros((n_pb, n_pa, n_bin), dtype=np.int64)
        o.pp_n = _ptr(res["pp_n"], _i64p)
        for key in ("pp_theta", "pp_gt", "pp_gx", "pp_w", "pp_k"):
            res[key] = np.zeros((n_pb, n_pa, n_bin))
            setattr(o, key, _ptr(res[key], _dp))
    rc = L.orc_correlate(C.byref(a), C.byref(o))
    if rc:
        raise RuntimeError(f"oracle failed with code {rc}")
    fin = finalize(res)
    res.update(fin)
    return res


def chord_edges_squared(edges_: np.ndarray) -> np.ndarray:
    """Squared chords 2 sin(e/2) of angular edges (the spherical bin test)."""
    e = _d(edges_)
    out = np.zeros_like(e)
    lib().orc_chord_edges_squared(_ptr(e, _dp), len(e) - 1, _ptr(out, _dp))
    return out


def sph_pair(lon1, lat1, lon2, lat2):
    """Per-pair great-circle angle and cos 2phi, sin 2phi of angles_sphere
    (src/utils.f90:618-639) in FP64."""
    a = [_d(np.atleast_1d(v)) for v in (lon1, lat1, lon2, lat2)]
    n = len(a[0])
    d, c2, s2 = np.zeros(n), np.zeros(n), np.zeros(n)
    lib().orc_sph_pair(n, *(_ptr(v, _dp) for v in a), _ptr(d, _dp), _ptr(c2, _dp), _ptr(s2, _dp))
    return d, c2, s2


def finalize(res: dict) -> dict:
    n_bin = len(res["theta"])
    mt, gt, gx = np.zeros(n_bin), np.zeros(n_bin), np.zeros(n_bin)
    lib().orc_finalize(n_bin, *[_ptr(_d(res[k]), _dp) for k in ("theta", "gt", "gx", "w", "k")],
                       _ptr(mt, _dp), _ptr(gt, _dp), _ptr(gx, _dp))
    return {"mean_theta": mt, "g_t": gt, "g_x": gx, "weights": np.asarray(res["w"], dtype=np.float64)}


def rp_bins(x_min, x_max, n_bin, spacing, ihs_theta=1.0):
    b = _RpBins()
    if lib().rp_create_bins(x_min, x_max, n_bin, spacing, ihs_theta, C.byref(b)):
        raise ValueError("ill-defined binning")
    return b


def rp_find_bin(x, x_min, x_max, n_bin, spacing, ihs_theta=1.0) -> np.ndarray:
    """Reference FP32 find_bin_index (1-based, -1 outside)."""
    b = rp_bins(x_min, x_max, n_bin, spacing, ihs_theta)
    x = np.ascontiguousarray(np.atleast_1d(x), dtype=np.float32)
    out = np.zeros(len(x), dtype=np.int32)
    lib().rp_find_bin_array(len(x), _ptr(x, _fp), C.byref(b), _ptr(out, _i32p))
    return out


def ref_port(kind: int, lens: dict, src: dict, x_min, x_max, n_bin, spacing, ihs_theta=1.0,
             n_threads: int = 1, use_tree: bool = True, marks: bool = False) -> dict:
    """The reference algorithm (kd-tree, FP32) for CPU timing."""
    keep = []

    def f(dct, key, n):
        v = dct.get(key)
        a = np.ascontiguousarray(v if v is not None else np.zeros(n), dtype=np.float32)
        keep.append(a)
        return _ptr(a, _fp)

    n_a, n_b = len(lens["x"]), len(src["x"])
    a = _RpArgs()
    a.kind, a.n_a, a.n_b = kind, n_a, n_b
    a.xa, a.ya, a.za, a.vala, a.wa = (f(lens, k, n_a) for k in ("x", "y", "z", "val", "w"))
    s1, s2 = ("e1", "e2") if kind == SHEAR else ("kappa_E", "kappa_B")
    a.xb, a.yb, a.zb, a.s1b, a.s2b, a.wb, a.mb = (f(src, k, n_b) for k in ("x", "y", "z", s1, s2, "w", "m"))
    a.bins = rp_bins(x_min, x_max, n_bin, spacing, ihs_theta)
    a.n_threads, a.use_tree = n_threads, 1 if use_tree else 0
    res = {k: np.zeros(n_bin, dtype=np.float32) for k in ("mean_theta", "g_t", "g_x", "weights")}
    res["n_pairs"] = np.zeros(n_bin, dtype=np.int64)
    o = _RpOut()
    for k in ("mean_theta", "g_t", "g_x", "weights"):
        setattr(o, k, _ptr(res[k], _fp))
    o.n_pairs = _ptr(res["n_pairs"], _i64p)
    if marks:
        for k in ("xi_e", "xi_b", "k_marks", "w_marks"):
            res[k] = np.zeros((n_a, n_bin), dtype=np.float32)
            setattr(o, k, _ptr(res[k], _fp))
        res["marks_weight"] = np.zeros(n_a, dtype=np.float32)
        o.marks_weight = _ptr(res["marks_weight"], _fp)
    rc = lib().rp_correlate(C.byref(a), C.byref(o))
    if rc:
        raise RuntimeError(f"ref_port failed with code {rc}")
    res["tree_seconds"] = o.tree_seconds
    return res


def have_reference() -> bool:
    return os.path.exists(REF_LIB_PATH)


def ref_lib():
    """The reference's own correlation code (``_ref/libref2pcf.so``)."""
    global _ref_lib
    if _ref_lib is None:
        if not have_reference():
            raise FileNotFoundError(f"{REF_LIB_PATH} is missing: run oracle/build_ref.sh where the reference is")
        L = C.CDLL(REF_LIB_PATH)
        ci, cf = C.c_int, C.c_float
        L.ref_shear_c.argtypes = [ci] + [_fp] * 5 + [ci] + [_fp] * 7 + [cf, cf, ci, ci, cf, ci, ci] + \
            [_fp] * 4 + [_i64p] + [_fp] * 5 + [ci]
        L.ref_kappa_c.argtypes = [ci] + [_fp] * 4 + [ci] + [_fp] * 5 + [cf, cf, ci, ci, cf, ci, ci] + \
            [_fp] * 4 + [_i64p] + [_fp] * 5 + [ci]
        L.ref_sph_pair_c.argtypes = [ci] + [_fp] * 7
        L.ref_covariance_c.argtypes = [ci, ci, _fp, _fp]
        L.ref_kmeans_c.argtypes = [ci, _fp, _fp, _fp, ci, ci, cf, ci, cf, cf, _i32p, _fp]
        L.ref_qsort_arg_c.argtypes = [ci, _fp, _i32p, ci]
        L.ref_bins_c.argtypes = [cf, cf, ci, ci, cf, ci, _fp, _fp, _fp, _i32p]
        L.ref_flush_c.argtypes = []
        _ref_lib = L
    return _ref_lib


def ref_covariance(x: np.ndarray) -> np.ndarray:
    """The reference's covariance_matrix (src/utils.f90:745-764, FP32) of
    samples x [n_samples][n_data]."""
    L = ref_lib()
    xf = np.asfortranarray(x, dtype=np.float32)
    n, d = xf.shape
    cov = np.zeros((d, d), dtype=np.float32, order="F")
    L.ref_covariance_c(n, d, _ptr(xf, _fp), _ptr(cov, _fp))
    return np.ascontiguousarray(cov)


def ref_sph_pair(lon1, lat1, lon2, lat2):
    """The reference's own distance_sphere and angles_sphere (FP32) per pair,
    with cos 2phi = 2 c^2 - 1 and sin 2phi = 2 c s as src/tpcf.f90:341-342."""
    L = ref_lib()
    a = [np.ascontiguousarray(np.atleast_1d(v), dtype=np.float32) for v in (lon1, lat1, lon2, lat2)]
    n = len(a[0])
    d, c2, s2 = (np.zeros(n, np.float32) for _ in range(3))
    L.ref_sph_pair_c(n, *(_ptr(v, _fp) for v in a), _ptr(d, _fp), _ptr(c2, _fp), _ptr(s2, _fp))
    return d, c2, s2


def reference(kind: int, lens: dict, src: dict, x_min, x_max, n_bin, spacing, ihs_theta=1.0,
              n_threads: int = 1, use_tree: bool = True, spherical: bool = False, quiet: bool = False) -> dict:
    """Runs the reference's tangential_shear_tpcf (kind 0) or scalar_kappa_tpcf
    (kind 1), src/tpcf.f90:190-545, on FP32 copies of the inputs.  Returns its
    outputs: mean_theta, g_t, g_x, weights, n_pairs and the per-lens marks
    (shape (n_A, n_bin)), plus the wall time of the call in ``seconds``."""
    import time
    L = ref_lib()
    keep = []

    def f(dct, key, n):
        v = dct.get(key)
        a = np.ascontiguousarray(v if v is not None else np.zeros(n), dtype=np.float32)
        keep.append(a)
        return _ptr(a, _fp)

    n_a, n_b = len(lens["x"]), len(src["x"])
    res = {k: np.zeros(n_bin, dtype=np.float32) for k in ("mean_theta", "g_t", "g_x", "weights")}
    res["n_pairs"] = np.zeros(n_bin, dtype=np.int64)
    for k in ("xi_e", "xi_b", "k_marks", "w_marks"):
        res[k] = np.zeros((max(n_a, 1), n_bin), dtype=np.float32)
    res["marks_weight"] = np.zeros(max(n_a, 1), dtype=np.float32)
    outs = [_ptr(res[k], _fp) for k in ("mean_theta", "g_t", "g_x", "weights")] + [_ptr(res["n_pairs"], _i64p)] + \
        [_ptr(res[k], _fp) for k in ("xi_e", "xi_b", "k_marks", "w_marks", "marks_weight")]
    common = [float(x_min), float(x_max), int(n_bin), int(spacing), float(ihs_theta), int(n_threads),
              1 if use_tree else 0]
    # quiet: the reference's messages (find_bin_index prints for every pair
    # FP32 rounding puts past x_max) go to /dev/null, not into the caller's
    # standard output
    saved = None
    if quiet:
        import sys
        sys.stdout.flush()
        saved = os.dup(1)
        null = os.open(os.devnull, os.O_WRONLY)
        os.dup2(null, 1)
        os.close(null)
    try:
        t0 = time.perf_counter()
        if kind == SHEAR:
            L.ref_shear_c(n_a, *(f(lens, k, n_a) for k in ("x", "y", "z", "val", "w")), n_b,
                          *(f(src, k, n_b) for k in ("x", "y", "z", "e1", "e2", "w", "m")), *common, *outs,
                          1 if spherical else 0)
        else:
            L.ref_kappa_c(n_a, *(f(lens, k, n_a) for k in ("x", "y", "val", "w")), n_b,
                          *(f(src, k, n_b) for k in ("x", "y", "kappa_E", "kappa_B", "w")), *common, *outs,
                          1 if spherical else 0)
        res["seconds"] = time.perf_counter() - t0
    finally:
        if saved is not None:
            L.ref_flush_c()
            os.dup2(saved, 1)
            os.close(saved)
    return res


def rect_block_labels(x, y, n_x_blocks: int, n_y_blocks: int, offset: int = 0) -> np.ndarray:
    """Block labels of one field in the rectangular_blocks scheme
    (src/tpcf.f90:1029-1049), 0-based: lin bins over the field's single
    precision bounding box (create_bins, src/utils.f90:139-230), find_bin_index
    through rp_find_bin (src/utils.f90:101-137), and the reference's index
    (iy - 1) * n_y_blocks + ix + offset minus one."""
    xs = np.asarray(x, dtype=np.float32)
    ys = np.asarray(y, dtype=np.float32)
    ix = rp_find_bin(xs, float(xs.min()), float(xs.max()), int(n_x_blocks), 0).astype(np.int64)
    iy = rp_find_bin(ys, float(ys.min()), float(ys.max()), int(n_y_blocks), 0).astype(np.int64)
    return (iy - 1) * int(n_y_blocks) + ix - 1 + int(offset)


def survey_block_sums(labels, w, xi_e, xi_b, k_marks, n_blocks: int) -> dict:
    """Block sums of survey_marked_point_block_bootstrap
    (src/tpcf.f90:1043-1069) from per-object marks [n][n_bin] and 0-based
    labels: sums in object order, then the empty blocks (weight 0) dropped
    and G_t, G_x set to 0 where a block's K is 0."""
    lab = np.asarray(labels, dtype=np.int64)
    nb = np.asarray(xi_e).shape[1]
    bw = np.zeros(n_blocks)
    bt = np.zeros((n_blocks, nb))
    bx = np.zeros((n_blocks, nb))
    bk = np.zeros((n_blocks, nb))
    for j in range(len(lab)):
        b = lab[j]
        bw[b] += w[j]
        bt[b] += xi_e[j]
        bx[b] += xi_b[j]
        bk[b] += k_marks[j]
    keep = bw > 0
    bt, bx, bk = bt[keep], bx[keep], bk[keep]
    bt[bk == 0] = 0.0
    bx[bk == 0] = 0.0
    return dict(g_t=bt, g_x=bx, k=bk, weight=bw[keep])


def bootstrap(blocks: dict, multiplicity: np.ndarray) -> dict:
    """The resampling of survey_marked_point_block_bootstrap
    (src/tpcf.f90:1165-1191) for given multiplicities M [n_boot][n_block]:
    g = (M . G) / (M . K) where M . K > 0, else 0; weights = M . w.  Sums in
    block order, FP64."""
    M = np.asarray(multiplicity, dtype=np.float64)
    gt, gx, k = (np.asarray(blocks[key], dtype=np.float64) for key in ("g_t", "g_x", "k"))
    st = np.zeros((M.shape[0], gt.shape[1]))
    sx = np.zeros_like(st)
    sk = np.zeros_like(st)
    for p in range(M.shape[1]):          # block order, as the resampling loop
        m = M[:, p:p + 1]
        st += m * gt[p]
        sx += m * gx[p]
        sk += m * k[p]
    with np.errstate(invalid="ignore", divide="ignore"):
        g_t = np.where(sk > 0, st / np.where(sk > 0, sk, 1), 0.0)
        g_x = np.where(sk > 0, sx / np.where(sk > 0, sk, 1), 0.0)
    w = blocks.get("weight")
    weights = M @ np.asarray(w, dtype=np.float64) if w is not None else np.zeros(M.shape[0])
    return dict(g_t=g_t, g_x=g_x, weights=weights)


def covariance(x: np.ndarray) -> np.ndarray:
    """covariance_matrix (src/utils.f90:745-764) in FP64: column means, then
    sum_i (x_ia - mean_a)(x_ib - mean_b) / (n - 1)."""
    x = np.asarray(x, dtype=np.float64)
    n = x.shape[0]
    if n <= 1:
        return np.zeros((x.shape[1], x.shape[1]))
    d = x - x.mean(axis=0)
    return d.T @ d / (n - 1)


def kmeans(x, y, w, k: int, max_iterations: int = 1000, reassignment_tol: float = 0.0, scaling: float = 0.0,
           jacobi: bool = False, spherical: bool = False, centroids=None, equal_size: bool = False,
           weight_tol: float = 0.0, block: int | None = None) -> dict:
    """partition_data (src/kmeans.f90:12-263) restated in FP64 (see
    orc_kmeans): 0-based labels, centroids [k][2], cluster weights, rounds,
    and with ``equal_size`` the post-processing's hand-overs.  ``block``: the
    equaliser's weights move after every block of that many points (1, the
    default: the reference's single thread); ``jacobi`` (block None): the
    previous round's weights for the whole round."""
    if block is None:
        block = 0 if jacobi else 1
    x, y, w = _d(x), _d(y), _d(w)
    n = len(x)
    lab = np.zeros(n, dtype=np.int32)
    cen = np.zeros((k, 2)) if centroids is None else np.array(centroids, dtype=np.float64)
    cw = np.zeros(k)
    moves, ppm = C.c_int64(0), C.c_int64(0)
    it = lib().orc_kmeans(n, _ptr(x, _dp), _ptr(y, _dp), _ptr(w, _dp), k, max_iterations, reassignment_tol, scaling,
                          int(block), 1 if spherical else 0, 0 if centroids is None else 1,
                          1 if equal_size else 0, float(weight_tol), _ptr(lab, _i32p), _ptr(cen, _dp), _ptr(cw, _dp),
                          C.byref(moves), C.byref(ppm))
    if it == -1:
        raise ValueError("k-means needs n >= 1 and k >= 2")
    if it < 0:
        raise MemoryError("oracle k-means: allocation failed")
    return dict(labels=lab, centroids=cen, cluster_weights=cw, iterations=it, last_moves=moves.value,
                pp_moves=ppm.value)


def qsort_arg(a) -> np.ndarray:
    """The oracle's restatement of qsort_arg (src/utils.f90:239-298): the
    0-based permutation sorting a in increasing order, ties in the
    reference's order."""
    a = _d(np.atleast_1d(a))
    order = np.zeros(len(a), dtype=np.int64)
    if lib().orc_qsort_arg(len(a), _ptr(a, _dp), order.ctypes.data_as(C.POINTER(C.c_int64))):
        raise MemoryError("oracle qsort_arg: allocation failed")
    return order


def ref_qsort_arg(a, reverse: bool = False) -> np.ndarray:
    """The reference's own qsort_arg (FP32 values), 0-based."""
    af = np.ascontiguousarray(a, dtype=np.float32)
    order = np.zeros(len(af), dtype=np.int32)
    ref_lib().ref_qsort_arg_c(len(af), _ptr(af, _fp), _ptr(order, _i32p), 1 if reverse else 0)
    return order.astype(np.int64) - 1


def ref_kmeans(x, y, w, k: int, max_iterations: int = 1000, reassignment_tol: float = 0.0, scaling: float = 0.0,
               equal_size: bool = False, weight_tol: float = 0.0) -> dict:
    """The reference's own partition_data (src/kmeans.f90:12-263, FP32, one
    thread, fixed initial centroids, Euclidean distance): 0-based labels."""
    L = ref_lib()
    xf, yf, wf = (np.ascontiguousarray(v, dtype=np.float32) for v in (x, y, w))
    n = len(xf)
    lab = np.zeros(n, dtype=np.int32)
    cen = np.zeros((2, k), dtype=np.float32)   # Fortran (k, 2) storage
    L.ref_kmeans_c(n, _ptr(xf, _fp), _ptr(yf, _fp), _ptr(wf, _fp), k, 1 if equal_size else 0, float(scaling),
                   int(max_iterations), float(reassignment_tol), float(weight_tol), _ptr(lab, _i32p), _ptr(cen, _fp))
    return dict(labels=lab - 1, centroids=cen.T.copy())
