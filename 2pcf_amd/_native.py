"""ctypes binding of the C-ABI declared in ``include/tpcf_gpu.h``.

The library is the in-tree ``lib2pcf_amd.so`` built by ``build.py``; there is
no CPU fallback: loading fails loudly when it is missing, and every
correlation call needs a GPU (``TPCF_ERR_NODEVICE`` otherwise).
"""
from __future__ import annotations

import ctypes as C
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "lib2pcf_amd.so")

MAX_BIN = 1024
GROUP_BINS = 63                              # bins per pair-kernel pass (tpcf_adjoint_*: the limit)
LIN, LOG, IHS, SQRT = 0, 1, 2, 3            # src/utils.f90:7
SHEAR, KAPPA = 0, 1

(OK, ERR_ARG, ERR_BINS, ERR_ALLOC, ERR_HIP, ERR_UNSUPPORTED, ERR_NONFINITE, ERR_NODEVICE, ERR_PATCH, ERR_GRID,
 ERR_MASK) = range(11)
STATUS_NAMES = {
    OK: "TPCF_OK", ERR_ARG: "TPCF_ERR_ARG", ERR_BINS: "TPCF_ERR_BINS", ERR_ALLOC: "TPCF_ERR_ALLOC",
    ERR_HIP: "TPCF_ERR_HIP", ERR_UNSUPPORTED: "TPCF_ERR_UNSUPPORTED", ERR_NONFINITE: "TPCF_ERR_NONFINITE",
    ERR_NODEVICE: "TPCF_ERR_NODEVICE", ERR_PATCH: "TPCF_ERR_PATCH", ERR_GRID: "TPCF_ERR_GRID",
    ERR_MASK: "TPCF_ERR_MASK",
}

FLAG_DEVICE_INPUTS = 1 << 0
FLAG_SPHERICAL = 1 << 1
FLAG_VERBOSE = 1 << 2
FLAG_NO_KD_TREE = 1 << 3
FLAG_BATCH_AUTOTUNE = 1 << 4
FLAG_BATCH_NO_GRAPH = 1 << 5
FLAG_DEVICE_MARKS = 1 << 6

OPT_CELL_SIZE, OPT_COUNTERS, OPT_CELL_TARGET, OPT_SPLIT = 1, 2, 3, 4

# every entry point the header declares
EXPORTS = (
    "tpcf_create_bins", "tpcf_find_bin_index", "tpcf_ctx_create", "tpcf_ctx_destroy", "tpcf_set_option",
    "tpcf_set_stream", "tpcf_get_stats", "tpcf_last_error",
    "tpcf_shear_f32", "tpcf_kappa_f32", "tpcf_shear_f64", "tpcf_kappa_f64",
    "tpcf_adjoint_shear_f32", "tpcf_adjoint_shear_f64", "tpcf_adjoint_kappa_f32", "tpcf_adjoint_kappa_f64",
    "tpcf_batch_create_f32", "tpcf_batch_create_f64", "tpcf_batch_run", "tpcf_batch_get_stats", "tpcf_batch_destroy",
    "tpcf_bootstrap", "tpcf_kmeans_f32", "tpcf_kmeans_f64",
    "tpcf_adjoint_record_shear_f32", "tpcf_adjoint_record_shear_f64", "tpcf_adjoint_record_kappa_f32",
    "tpcf_adjoint_record_kappa_f64", "tpcf_adjoint_replay", "tpcf_tape_destroy", "tpcf_shear_f32_noalloc",
    "tpcf_rect_blocks_f32", "tpcf_rect_blocks_f64",
)


class TpcfError(RuntimeError):
    def __init__(self, code: int, msg: str = ""):
        self.code = code
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {msg}" if msg else STATUS_NAMES.get(code, str(code)))


class Bins(C.Structure):
    """``tpcf_bins`` = type(binning), src/utils.f90:18-24, plus FP64 edges."""
    _fields_ = [
        ("x_min", C.c_float), ("x_max", C.c_float), ("bin_size", C.c_float),
        ("n_bin", C.c_int32), ("spacing", C.c_int32), ("ihs_theta", C.c_float),
        ("edges", C.c_double * (MAX_BIN + 1)), ("bin_centers", C.c_float * MAX_BIN),
    ]


class Out(C.Structure):
    _fields_ = [
        ("real_bytes", C.c_int32),
        ("theta", C.c_void_p), ("mean_theta", C.c_void_p), ("g_t", C.c_void_p), ("g_x", C.c_void_p),
        ("weights", C.c_void_p), ("n_pairs", C.c_void_p),
        ("sum_theta", C.c_void_p), ("sum_gt", C.c_void_p), ("sum_gx", C.c_void_p), ("sum_w", C.c_void_p),
        ("sum_k", C.c_void_p),
        ("xi_E_marks", C.c_void_p), ("xi_B_marks", C.c_void_p), ("K_marks", C.c_void_p), ("w_marks", C.c_void_p),
        ("marks_weight", C.c_void_p),
        ("pp_n", C.c_void_p), ("pp_theta", C.c_void_p), ("pp_gt", C.c_void_p), ("pp_gx", C.c_void_p),
        ("pp_w", C.c_void_p), ("pp_k", C.c_void_p),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("ms_total", C.c_double), ("ms_upload", C.c_double), ("ms_prepare", C.c_double), ("ms_pairs", C.c_double),
        ("ms_finish", C.c_double), ("n_pairs", C.c_int64), ("n_lens_used", C.c_int64),
        ("n_source_used", C.c_int64), ("n_chunks", C.c_int64), ("n_segments", C.c_int64),
        ("cell_size", C.c_double), ("seg_visits", C.c_int64 * 3), ("pair_visits", C.c_int64 * 3),
        ("n_items", C.c_int64),
    ]

    def as_dict(self) -> dict:
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["seg_visits"] = list(self.seg_visits)
        d["pair_visits"] = list(self.pair_visits)
        return d


class Catalog(C.Structure):
    """``tpcf_catalog``: one catalogue of a batch (device pointers)."""
    _fields_ = [
        ("kind", C.c_int32), ("n_patch_A", C.c_int32), ("n_patch_B", C.c_int32),
        ("n_A", C.c_int64), ("n_B", C.c_int64),
        ("x_A", C.c_void_p), ("y_A", C.c_void_p), ("z_A", C.c_void_p), ("val_A", C.c_void_p), ("w_A", C.c_void_p),
        ("x_B", C.c_void_p), ("y_B", C.c_void_p), ("z_B", C.c_void_p), ("s1_B", C.c_void_p), ("s2_B", C.c_void_p),
        ("w_B", C.c_void_p), ("m_B", C.c_void_p),
        ("patch_A", C.c_void_p), ("patch_B", C.c_void_p),
        ("out", C.POINTER(Out)),
    ]


class BatchStats(C.Structure):
    _fields_ = [
        ("ms_plan", C.c_double), ("ms_run", C.c_double), ("cell_target", C.c_double), ("split", C.c_double),
        ("tune_ms", C.c_double * 5), ("n_pairs", C.c_int64), ("n_items", C.c_int64), ("graph_nodes", C.c_int64),
        ("runs", C.c_int64), ("n_catalogs", C.c_int32), ("autotuned", C.c_int32),
    ]

    def as_dict(self) -> dict:
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["tune_ms"] = list(self.tune_ms)
        return d


class KmeansOpts(C.Structure):
    _fields_ = [("k", C.c_int32), ("max_iterations", C.c_int32), ("init", C.c_int32), ("equal_size", C.c_int32),
                ("reassignment_tol", C.c_double), ("weight_equalizer_scaling", C.c_double),
                ("weight_tol", C.c_double)]


class Blocks(C.Structure):
    _fields_ = [("n_block", C.c_int32), ("n_bin", C.c_int32), ("g_t", C.c_void_p), ("g_x", C.c_void_p),
                ("k", C.c_void_p), ("weight", C.c_void_p)]


_lib = None
_vp = C.c_void_p
_i64 = C.c_int64
_i32 = C.c_int32


def lib():
    """Load the in-tree library (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build() (or python 2pcf_amd/build.py)")
    # PyTorch-ROCm ships its own HIP runtime (libamdhip64.so) and makes its
    # symbols global.  Loaded first, it is also the runtime this library binds
    # to; loaded after this library (which needs the system's
    # libamdhip64.so.7) it would be a second HIP runtime in the process, and
    # torch's device initialisation then fails ("No HIP GPUs are available").
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    L.tpcf_create_bins.argtypes = [C.c_float, C.c_float, _i32, _i32, C.c_float, C.POINTER(Bins)]
    L.tpcf_create_bins.restype = C.c_int
    L.tpcf_find_bin_index.argtypes = [C.c_double, C.POINTER(Bins)]
    L.tpcf_find_bin_index.restype = _i32
    L.tpcf_ctx_create.argtypes = [_i32, C.POINTER(_vp)]
    L.tpcf_ctx_create.restype = C.c_int
    L.tpcf_ctx_destroy.argtypes = [_vp]
    L.tpcf_ctx_destroy.restype = C.c_int
    L.tpcf_set_option.argtypes = [_vp, _i32, C.c_double]
    L.tpcf_set_option.restype = C.c_int
    L.tpcf_set_stream.argtypes = [_vp, _vp]
    L.tpcf_set_stream.restype = C.c_int
    L.tpcf_get_stats.argtypes = [_vp, C.POINTER(Stats)]
    L.tpcf_get_stats.restype = C.c_int
    L.tpcf_last_error.argtypes = [_vp]
    L.tpcf_last_error.restype = C.c_char_p
    shear_args = [_vp, _i64] + [_vp] * 5 + [_i64] + [_vp] * 7 + [C.POINTER(Bins), C.c_uint32, _vp, _vp, _i32, _i32,
                                                                  C.POINTER(Out)]
    kappa_args = [_vp, _i64] + [_vp] * 4 + [_i64] + [_vp] * 5 + [C.POINTER(Bins), C.c_uint32, _vp, _vp, _i32, _i32,
                                                                 C.POINTER(Out)]
    for name, args in (("tpcf_shear_f32", shear_args), ("tpcf_shear_f64", shear_args),
                       ("tpcf_kappa_f32", kappa_args), ("tpcf_kappa_f64", kappa_args),
                       ("tpcf_shear_f32_noalloc", shear_args)):
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    adj_shear = shear_args[:-7] + [C.POINTER(Bins), C.c_uint32, _vp, _vp] + [_vp] * 6
    adj_kappa = kappa_args[:-7] + [C.POINTER(Bins), C.c_uint32, _vp, _vp] + [_vp] * 5
    for name, args in (("tpcf_adjoint_shear_f32", adj_shear), ("tpcf_adjoint_shear_f64", adj_shear),
                       ("tpcf_adjoint_kappa_f32", adj_kappa), ("tpcf_adjoint_kappa_f64", adj_kappa)):
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    rec_shear = shear_args[:-7] + [C.POINTER(Bins), C.c_uint32, _vp, _vp, C.POINTER(_vp)]
    rec_kappa = kappa_args[:-7] + [C.POINTER(Bins), C.c_uint32, _vp, _vp, C.POINTER(_vp)]
    for name, args in (("tpcf_adjoint_record_shear_f32", rec_shear), ("tpcf_adjoint_record_shear_f64", rec_shear),
                       ("tpcf_adjoint_record_kappa_f32", rec_kappa), ("tpcf_adjoint_record_kappa_f64", rec_kappa)):
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    L.tpcf_adjoint_replay.argtypes = [_vp, _vp, _vp, _vp] + [_vp] * 6
    L.tpcf_adjoint_replay.restype = C.c_int
    L.tpcf_tape_destroy.argtypes = [_vp]
    L.tpcf_tape_destroy.restype = C.c_int
    for name in ("tpcf_batch_create_f32", "tpcf_batch_create_f64"):
        f = getattr(L, name)
        f.argtypes = [_vp, _i32, C.POINTER(Catalog), C.POINTER(Bins), C.c_uint32, C.POINTER(_vp)]
        f.restype = C.c_int
    L.tpcf_batch_run.argtypes = [_vp]
    L.tpcf_batch_run.restype = C.c_int
    L.tpcf_batch_get_stats.argtypes = [_vp, C.POINTER(BatchStats)]
    L.tpcf_batch_get_stats.restype = C.c_int
    L.tpcf_batch_destroy.argtypes = [_vp]
    L.tpcf_batch_destroy.restype = C.c_int
    L.tpcf_bootstrap.argtypes = [_vp, C.POINTER(Blocks), _i32, _vp, C.c_uint64, C.c_double, C.c_uint32] + [_vp] * 6
    L.tpcf_bootstrap.restype = C.c_int
    for name in ("tpcf_kmeans_f32", "tpcf_kmeans_f64"):
        f = getattr(L, name)
        f.argtypes = [_vp, _i64, _vp, _vp, _vp, C.POINTER(KmeansOpts), C.c_uint32, _vp, _vp, _vp,
                      C.POINTER(_i32), C.POINTER(_i64)]
        f.restype = C.c_int
    for name in ("tpcf_rect_blocks_f32", "tpcf_rect_blocks_f64"):
        f = getattr(L, name)
        f.argtypes = [_vp, _i64, _vp, _vp, _i32, _i32, _i32, C.c_uint32, _vp]
        f.restype = C.c_int
    _lib = L
    return L
