"""Host-side mirror of the reference's ``tpcf`` / ``utils`` module interface for
the correlation call, running on the gfx950 library through its C-ABI.

Reference (tilmantroester/2pcf):
  * ``binning`` / ``create_bins`` / ``find_bin_index``   src/utils.f90:18-24, :139-230, :101-137
  * ``shape_catalog``, ``lens_catalog``, ``kappa_map``, ``lensing_survey_field``,
    ``shear_2pcf_results``                                 src/tpcf.f90:17-43
  * ``shear_2pcf`` -> tangential_shear_tpcf             src/tpcf.f90:134-160, :190-378
  * ``kappa_2pcf`` -> scalar_kappa_tpcf                 src/tpcf.f90:162-188, :380-545
  * field stacking                                       src/main.f90:678-695

Differences from the reference, kept on purpose: marks can be switched off
(``marks=False``; the reference always computes them because
``calculate_marks`` is never set, src/tpcf.f90:146); per-patch-pair sums and
raw FP64 sums are returned as well; errors raise ``TpcfError`` instead of
stopping the program.  ``num_threads`` and ``use_kd_tree`` are accepted and
have no effect (the pair set is the reference's "every pair within x_max").
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _native as N
from ._native import IHS, KAPPA, LIN, LOG, SHEAR, SQRT, TpcfError  # noqa: F401  (re-exported)

lin_binning, log_binning, ihs_binning, sqrt_binning = LIN, LOG, IHS, SQRT


# ---------------------------------------------------------------- binning ----
@dataclass
class Binning:
    """type(binning), src/utils.f90:18-24; ``bin_edges`` are the FP64 edges the
    library bins with (squared-edge comparison in FP64)."""
    x_min: float
    x_max: float
    bin_size: float
    n_bin: int
    spacing: int
    ihs_theta: float
    bin_edges: np.ndarray
    bin_centers: np.ndarray
    c: N.Bins = field(repr=False, default=None)


def create_bins(x_min: float, x_max: float, n_bin: int, spacing: int = LOG, p: float = 1.0) -> Binning:
    """create_bins(x_min=, x_max=, n_bin=, p=, spacing=), src/utils.f90:139-230."""
    b = N.Bins()
    rc = N.lib().tpcf_create_bins(x_min, x_max, n_bin, spacing, p, C.byref(b))
    if rc:
        raise TpcfError(rc, "ill-defined binning")
    n = b.n_bin
    return Binning(x_min=b.x_min, x_max=b.x_max, bin_size=b.bin_size, n_bin=n, spacing=b.spacing,
                   ihs_theta=b.ihs_theta, bin_edges=np.array(b.edges[: n + 1]),
                   bin_centers=np.array(b.bin_centers[:n], dtype=np.float32), c=b)


def find_bin_index(x: float, bins: Binning) -> int:
    """find_bin_index, src/utils.f90:101-137: 1-based bin, -1 outside."""
    k = N.lib().tpcf_find_bin_index(float(x), C.byref(bins.c))
    return k + 1 if k >= 0 else -1


# ---------------------------------------------------------------- catalogues ----
@dataclass
class ShapeCatalog:       # src/tpcf.f90:24-27
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    e1: np.ndarray
    e2: np.ndarray
    m: np.ndarray
    w: np.ndarray

    @property
    def n(self) -> int:
        return len(self.x)


@dataclass
class KappaMap:           # src/tpcf.f90:29-32
    x: np.ndarray
    y: np.ndarray
    kappa_E: np.ndarray
    kappa_B: np.ndarray
    w: np.ndarray

    @property
    def n(self) -> int:
        return len(self.x)


@dataclass
class LensCatalog:        # src/tpcf.f90:34-37
    x: np.ndarray
    y: np.ndarray
    val: np.ndarray
    w: np.ndarray
    z: Optional[np.ndarray] = None

    @property
    def n(self) -> int:
        return len(self.x)


@dataclass
class LensingSurveyField:  # src/tpcf.f90:39-43
    lenses: LensCatalog
    shapes: Optional[ShapeCatalog] = None
    convergence: Optional[KappaMap] = None


@dataclass
class Shear2pcfResults:   # src/tpcf.f90:17-22 plus raw and per-patch-pair sums
    n_bin: int
    theta: np.ndarray
    mean_theta: np.ndarray
    g_t: np.ndarray
    g_x: np.ndarray
    weights: np.ndarray
    n_pairs: np.ndarray
    sum_theta: np.ndarray
    sum_gt: np.ndarray
    sum_gx: np.ndarray
    sum_w: np.ndarray
    sum_k: np.ndarray
    # per-lens marks, shape (n_bin, n_A) in Fortran order like real(n_bin, n_A)
    xi_E_marks: Optional[np.ndarray] = None
    xi_B_marks: Optional[np.ndarray] = None
    K_marks: Optional[np.ndarray] = None
    w_marks: Optional[np.ndarray] = None
    marks_weight: Optional[np.ndarray] = None
    # per-patch-pair raw sums, shape (P_B, P_A, n_bin) == real(n_bin, P_A, P_B)
    pp_n: Optional[np.ndarray] = None
    pp_theta: Optional[np.ndarray] = None
    pp_gt: Optional[np.ndarray] = None
    pp_gx: Optional[np.ndarray] = None
    pp_w: Optional[np.ndarray] = None
    pp_k: Optional[np.ndarray] = None
    stats: Optional[dict] = None


# ---------------------------------------------------------------- context ----
class Context:
    """Owns a ``tpcf_ctx`` (device, stream, reusable device buffers)."""

    def __init__(self, device: int = -1):
        self.device = device   # -1: the device current at creation
        self._h = C.c_void_p()
        rc = N.lib().tpcf_ctx_create(device, C.byref(self._h))
        if rc:
            raise TpcfError(rc, "tpcf_ctx_create failed")

    def close(self) -> None:
        if self._h:
            N.lib().tpcf_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    def set_option(self, option: int, value: float) -> None:
        rc = N.lib().tpcf_set_option(self._h, option, float(value))
        if rc:
            raise TpcfError(rc, self.last_error())

    def set_stream(self, stream_handle: int | None) -> None:
        N.lib().tpcf_set_stream(self._h, C.c_void_p(stream_handle) if stream_handle else None)

    def last_error(self) -> str:
        s = N.lib().tpcf_last_error(self._h)
        return s.decode() if s else ""

    def stats(self) -> dict:
        st = N.Stats()
        N.lib().tpcf_get_stats(self._h, C.byref(st))
        return st.as_dict()


_default_ctx: Optional[Context] = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context()
    return _default_ctx


# ---------------------------------------------------------------- core call ----
def _host(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def correlate(kind: int, lens: dict, src: dict, bins: Binning, *, precision: str = "f32", marks: bool = False,
              patch_A=None, patch_B=None, n_patch_A: int = 1, n_patch_B: int = 1, real_bytes: int = 4,
              device_inputs: bool = False, verbose: bool = False, ctx: Optional[Context] = None,
              n_A: Optional[int] = None, n_B: Optional[int] = None, spherical: bool = False,
              marks_device: bool = False, allocation_free: bool = False) -> Shear2pcfResults:
    """One call of the C-ABI.  ``lens``: x, y, val, w (+ z for shear);
    ``src``: x, y, w + (z, e1, e2, m) for shear or (kappa_E, kappa_B) for kappa.
    With ``device_inputs`` the values are device addresses (ints) of arrays of
    the given precision and ``n_A``/``n_B`` give the sizes.  With ``spherical``
    x, y are longitude and latitude and the bins angles, all in radians.  With
    ``marks`` and ``marks_device`` the marks are torch tensors on the current
    GPU (TPCF_DEVICE_MARKS: no copy to the host).  ``allocation_free``
    (shear, FP32) calls tpcf_shear_f32_noalloc, which carves its device
    buffers from the library's fixed scratch array instead of allocating."""
    ctx = ctx or default_context()
    L = N.lib()
    dt = np.float32 if precision == "f32" else np.float64
    keep = []

    def ptr(v, dtype=dt):
        if v is None:
            return None
        if device_inputs:
            return C.c_void_p(int(v))
        a = _host(v, dtype)
        keep.append(a)
        return a.ctypes.data_as(C.c_void_p)

    if not device_inputs:
        n_A = len(lens["x"])
        n_B = len(src["x"])
    nb = bins.n_bin
    rt = np.float32 if real_bytes == 4 else np.float64
    res = {k: np.zeros(nb, dtype=rt) for k in ("theta", "mean_theta", "g_t", "g_x", "weights")}
    res["n_pairs"] = np.zeros(nb, dtype=np.int64)
    for k in ("sum_theta", "sum_gt", "sum_gx", "sum_w", "sum_k"):
        res[k] = np.zeros(nb, dtype=np.float64)
    o = N.Out()
    o.real_bytes = real_bytes
    for k, v in res.items():
        setattr(o, k, v.ctypes.data_as(C.c_void_p))
    if marks and marks_device:
        import torch
        tt = torch.float32 if real_bytes == 4 else torch.float64
        dev = torch.device("cuda", ctx.device) if ctx.device >= 0 else torch.device("cuda")
        for k in ("xi_E_marks", "xi_B_marks", "K_marks", "w_marks"):
            res[k] = torch.empty((n_A, nb), dtype=tt, device=dev)
            setattr(o, k, C.c_void_p(res[k].data_ptr()))
        res["marks_weight"] = torch.empty(n_A, dtype=tt, device=dev)
        o.marks_weight = C.c_void_p(res["marks_weight"].data_ptr())
        # the library writes them on its own stream: torch's pending work on
        # this memory (a reuse by the caching allocator) ends first
        torch.cuda.current_stream(dev).synchronize()
    elif marks:
        for k in ("xi_E_marks", "xi_B_marks", "K_marks", "w_marks"):
            res[k] = np.zeros((n_A, nb), dtype=rt)
            setattr(o, k, res[k].ctypes.data_as(C.c_void_p))
        res["marks_weight"] = np.zeros(n_A, dtype=rt)
        o.marks_weight = res["marks_weight"].ctypes.data_as(C.c_void_p)
    patches = (patch_A is not None) or (patch_B is not None)
    if patches:
        P_A = n_patch_A if patch_A is not None else 1
        P_B = n_patch_B if patch_B is not None else 1
        res["pp_n"] = np.zeros((P_B, P_A, nb), dtype=np.int64)
        o.pp_n = res["pp_n"].ctypes.data_as(C.c_void_p)
        for k in ("pp_theta", "pp_gt", "pp_gx", "pp_w", "pp_k"):
            res[k] = np.zeros((P_B, P_A, nb), dtype=np.float64)
            setattr(o, k, res[k].ctypes.data_as(C.c_void_p))
    flags = (N.FLAG_DEVICE_INPUTS if device_inputs else 0) | (N.FLAG_VERBOSE if verbose else 0) | \
        (N.FLAG_SPHERICAL if spherical else 0) | (N.FLAG_DEVICE_MARKS if marks and marks_device else 0)
    pa = ptr(patch_A, np.int32) if patch_A is not None else None
    pb = ptr(patch_B, np.int32) if patch_B is not None else None
    if kind == SHEAR:
        fn = L.tpcf_shear_f32 if precision == "f32" else L.tpcf_shear_f64
        if allocation_free:
            if precision != "f32":
                raise ValueError("the allocation-free entry point is the FP32 shear call")
            fn = L.tpcf_shear_f32_noalloc
        rc = fn(ctx.handle, n_A, ptr(lens["x"]), ptr(lens["y"]), ptr(lens["z"]), ptr(lens["val"]), ptr(lens["w"]),
                n_B, ptr(src["x"]), ptr(src["y"]), ptr(src["z"]), ptr(src["e1"]), ptr(src["e2"]), ptr(src["w"]),
                ptr(src["m"]), C.byref(bins.c), flags, pa, pb, n_patch_A, n_patch_B, C.byref(o))
    else:
        if allocation_free:
            raise ValueError("the allocation-free entry point is the FP32 shear call")
        fn = L.tpcf_kappa_f32 if precision == "f32" else L.tpcf_kappa_f64
        rc = fn(ctx.handle, n_A, ptr(lens["x"]), ptr(lens["y"]), ptr(lens["val"]), ptr(lens["w"]),
                n_B, ptr(src["x"]), ptr(src["y"]), ptr(src["kappa_E"]), ptr(src["kappa_B"]), ptr(src["w"]),
                C.byref(bins.c), flags, pa, pb, n_patch_A, n_patch_B, C.byref(o))
    if rc:
        raise TpcfError(rc, ctx.last_error())
    out = Shear2pcfResults(n_bin=nb, stats=ctx.stats(), **{k: v for k, v in res.items() if not k.endswith("marks")
                                                             and k != "marks_weight"})
    if marks:
        for k in ("xi_E_marks", "xi_B_marks", "K_marks", "w_marks"):
            setattr(out, k, res[k].T)   # (n_bin, n_A), Fortran order, no copy
        out.marks_weight = res["marks_weight"]
    return out


def _precision_of(arr) -> str:
    return "f64" if np.asarray(arr).dtype == np.float64 else "f32"


def shear_2pcf(field: LensingSurveyField, bins: Binning, results=None, spherical_coords: bool = False,
               num_threads: int = 1, use_kd_tree: bool = True, verbose: bool = False, *, marks: bool = True,
               precision: Optional[str] = None, patch_A=None, patch_B=None, n_patch_A: int = 1, n_patch_B: int = 1,
               ctx: Optional[Context] = None) -> Shear2pcfResults:
    """shear_2pcf(field, bins, results, spherical_coords, num_threads,
    use_kd_tree, verbose), src/tpcf.f90:134-160.  Returns the results object
    (the reference fills its ``results`` argument).  With spherical_coords
    positions are longitude / latitude and the bins angles, in radians (the
    reference's driver converts degrees before the call, src/main.f90:427-431,
    :597-603)."""
    L, S = field.lenses, field.shapes
    prec = precision or _precision_of(L.x)
    lens = dict(x=L.x, y=L.y, z=L.z if L.z is not None else np.zeros(L.n), val=L.val, w=L.w)
    src = dict(x=S.x, y=S.y, z=S.z, e1=S.e1, e2=S.e2, w=S.w, m=S.m)
    return correlate(SHEAR, lens, src, bins, precision=prec, marks=marks, patch_A=patch_A, patch_B=patch_B,
                     n_patch_A=n_patch_A, n_patch_B=n_patch_B, real_bytes=4 if prec == "f32" else 8,
                     verbose=verbose, ctx=ctx, spherical=spherical_coords)


def kappa_2pcf(field: LensingSurveyField, bins: Binning, results=None, spherical_coords: bool = False,
               num_threads: int = 1, use_kd_tree: bool = True, verbose: bool = False, *, marks: bool = True,
               precision: Optional[str] = None, patch_A=None, patch_B=None, n_patch_A: int = 1, n_patch_B: int = 1,
               ctx: Optional[Context] = None) -> Shear2pcfResults:
    """kappa_2pcf(field, bins, results, spherical_coords, num_threads,
    use_kd_tree, verbose), src/tpcf.f90:162-188."""
    L, K = field.lenses, field.convergence
    prec = precision or _precision_of(L.x)
    lens = dict(x=L.x, y=L.y, val=L.val, w=L.w)
    src = dict(x=K.x, y=K.y, kappa_E=K.kappa_E, kappa_B=K.kappa_B, w=K.w)
    return correlate(KAPPA, lens, src, bins, precision=prec, marks=marks, patch_A=patch_A, patch_B=patch_B,
                     n_patch_A=n_patch_A, n_patch_B=n_patch_B, real_bytes=4 if prec == "f32" else 8,
                     verbose=verbose, ctx=ctx, spherical=spherical_coords)


def stack_fields(results: list[Shear2pcfResults]) -> dict:
    """Weight-stacking of per-field results, src/main.f90:678-695:
    g = sum_f W_f g_f / sum_f W_f, n_pairs = sum_f N_f, mean_theta likewise."""
    W = sum(np.asarray(r.weights, dtype=np.float64) for r in results)
    with np.errstate(invalid="ignore", divide="ignore"):
        g_t = sum(np.asarray(r.weights, np.float64) * np.asarray(r.g_t, np.float64) for r in results) / W
        g_x = sum(np.asarray(r.weights, np.float64) * np.asarray(r.g_x, np.float64) for r in results) / W
        mt = sum(np.asarray(r.weights, np.float64) * np.asarray(r.mean_theta, np.float64) for r in results) / W
    return dict(theta=results[0].theta, g_t=g_t, g_x=g_x, weights=W,
                n_pairs=sum(np.asarray(r.n_pairs) for r in results), mean_theta=mt)


# ---------------------------------------------------------------- adjoint ----
_GRAD_KEYS = {SHEAR: (("val", "w"), ("e1", "e2", "w", "m")), KAPPA: (("val", "w"), ("kappa_E", "kappa_B", "w"))}


def adjoint(kind: int, lens: dict, src: dict, bins: Binning, gbar1, gbar2, *, precision: str = "f32",
            device_inputs: bool = False, grads: Optional[dict] = None, ctx: Optional[Context] = None,
            n_A: Optional[int] = None, n_B: Optional[int] = None, spherical: bool = False) -> dict:
    """Gradient of sum_b (gbar1_b g_t,b + gbar2_b g_x,b) (shear; xi_E / xi_B
    for kappa) with respect to the lens values and weights and the source
    components and weights: {"lens": {"val", "w"}, "src": {"e1", "e2", "w",
    "m"} or {"kappa_E", "kappa_B", "w"}} as FP64 arrays.  With
    ``device_inputs`` the inputs are device addresses and ``grads`` maps the
    same keys to device addresses of FP64 arrays to fill (missing keys are not
    computed); the return value is then empty."""
    ctx = ctx or default_context()
    L = N.lib()
    dt = np.float32 if precision == "f32" else np.float64
    keep = []

    def ptr(v, dtype=dt):
        if v is None:
            return None
        if device_inputs:
            return C.c_void_p(int(v))
        a = _host(v, dtype)
        keep.append(a)
        return a.ctypes.data_as(C.c_void_p)

    if not device_inputs:
        n_A, n_B = len(lens["x"]), len(src["x"])
    g1 = np.ascontiguousarray(gbar1, dtype=np.float64)
    g2 = np.ascontiguousarray(gbar2, dtype=np.float64)
    if len(g1) != bins.n_bin or len(g2) != bins.n_bin:
        raise ValueError("upstream gradients must have n_bin entries")
    lk, sk = _GRAD_KEYS[kind]
    out = {"lens": {}, "src": {}}
    gptr = []
    for side, keys, n in (("lens", lk, n_A), ("src", sk, n_B)):
        for k in keys:
            if device_inputs:
                addr = (grads or {}).get(side, {}).get(k)
                gptr.append(C.c_void_p(int(addr)) if addr is not None else None)
            else:
                a = np.zeros(n, dtype=np.float64)
                out[side][k] = a
                gptr.append(a.ctypes.data_as(C.c_void_p))
    flags = (N.FLAG_DEVICE_INPUTS if device_inputs else 0) | (N.FLAG_SPHERICAL if spherical else 0)
    pc = lambda a: g1.ctypes.data_as(C.c_void_p) if a is g1 else g2.ctypes.data_as(C.c_void_p)  # noqa: E731
    if kind == SHEAR:
        fn = L.tpcf_adjoint_shear_f32 if precision == "f32" else L.tpcf_adjoint_shear_f64
        rc = fn(ctx.handle, n_A, ptr(lens["x"]), ptr(lens["y"]), ptr(lens["z"]), ptr(lens["val"]), ptr(lens["w"]),
                n_B, ptr(src["x"]), ptr(src["y"]), ptr(src["z"]), ptr(src["e1"]), ptr(src["e2"]), ptr(src["w"]),
                ptr(src["m"]), C.byref(bins.c), flags, pc(g1), pc(g2), *gptr)
    else:
        fn = L.tpcf_adjoint_kappa_f32 if precision == "f32" else L.tpcf_adjoint_kappa_f64
        rc = fn(ctx.handle, n_A, ptr(lens["x"]), ptr(lens["y"]), ptr(lens["val"]), ptr(lens["w"]),
                n_B, ptr(src["x"]), ptr(src["y"]), ptr(src["kappa_E"]), ptr(src["kappa_B"]), ptr(src["w"]),
                C.byref(bins.c), flags, pc(g1), pc(g2), *gptr)
    if rc:
        raise TpcfError(rc, ctx.last_error())
    return out if not device_inputs else {}


class Tape:
    """A recorded lens-side pass (``tpcf_tape``): ``adjoint_replay`` gives the
    gradient for any upstream gradient with the source-side pass only.  Holds
    device memory (4 n_bin doubles per used lens) until closed or collected."""

    def __init__(self, handle: C.c_void_p, ctx: Context, kind: int, n_A: int, n_B: int, n_bin: int):
        self._h, self.ctx, self.kind, self.n_A, self.n_B, self.n_bin = handle, ctx, kind, n_A, n_B, n_bin

    def close(self) -> None:
        if self._h:
            N.lib().tpcf_tape_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self) -> C.c_void_p:
        return self._h


def adjoint_record(kind: int, lens: dict, src: dict, bins: Binning, *, precision: str = "f32",
                   ctx: Optional[Context] = None, n_A: int, n_B: int, spherical: bool = False):
    """Forward outputs (g_t, g_x or xi_E, xi_B: FP64 arrays [n_bin]) and a
    :class:`Tape` for ``adjoint_replay``.  ``lens`` / ``src`` hold device
    addresses (ints) of arrays of the given precision, which must stay valid
    and unchanged while the tape is replayed.  ``spherical``: x, y are
    longitude and latitude and the bin edges angles, all in radians."""
    ctx = ctx or default_context()
    L = N.lib()
    p = lambda v: C.c_void_p(int(v)) if v is not None else None   # noqa: E731
    g1 = np.zeros(bins.n_bin, np.float64)
    g2 = np.zeros(bins.n_bin, np.float64)
    h = C.c_void_p()
    flags = N.FLAG_DEVICE_INPUTS | (N.FLAG_SPHERICAL if spherical else 0)
    if kind == SHEAR:
        fn = L.tpcf_adjoint_record_shear_f32 if precision == "f32" else L.tpcf_adjoint_record_shear_f64
        rc = fn(ctx.handle, n_A, p(lens["x"]), p(lens["y"]), p(lens["z"]), p(lens["val"]), p(lens["w"]),
                n_B, p(src["x"]), p(src["y"]), p(src["z"]), p(src["e1"]), p(src["e2"]), p(src["w"]), p(src["m"]),
                C.byref(bins.c), flags, g1.ctypes.data_as(C.c_void_p),
                g2.ctypes.data_as(C.c_void_p), C.byref(h))
    else:
        fn = L.tpcf_adjoint_record_kappa_f32 if precision == "f32" else L.tpcf_adjoint_record_kappa_f64
        rc = fn(ctx.handle, n_A, p(lens["x"]), p(lens["y"]), p(lens["val"]), p(lens["w"]),
                n_B, p(src["x"]), p(src["y"]), p(src["kappa_E"]), p(src["kappa_B"]), p(src["w"]),
                C.byref(bins.c), flags, g1.ctypes.data_as(C.c_void_p),
                g2.ctypes.data_as(C.c_void_p), C.byref(h))
    if rc:
        raise TpcfError(rc, ctx.last_error())
    return g1, g2, Tape(h, ctx, kind, n_A, n_B, bins.n_bin)


def adjoint_replay(tape: Tape, gbar1, gbar2, grads: dict) -> None:
    """Fill ``grads`` ({"lens": {"val", "w"}, "src": {...}} -> device
    addresses of FP64 arrays; missing keys are not computed) with the gradient
    of sum_b (gbar1_b g_1,b + gbar2_b g_2,b) for the recorded pass."""
    L = N.lib()
    g1 = np.ascontiguousarray(gbar1, dtype=np.float64)
    g2 = np.ascontiguousarray(gbar2, dtype=np.float64)
    if len(g1) != tape.n_bin or len(g2) != tape.n_bin:
        raise ValueError("upstream gradients must have n_bin entries")
    lk, sk = _GRAD_KEYS[tape.kind]
    gptr = []
    for side, keys in (("lens", lk), ("src", sk)):
        for k in keys:
            addr = (grads or {}).get(side, {}).get(k)
            gptr.append(C.c_void_p(int(addr)) if addr is not None else None)
    while len(gptr) < 6:
        gptr.append(None)
    rc = L.tpcf_adjoint_replay(tape.ctx.handle, tape.handle, g1.ctypes.data_as(C.c_void_p),
                               g2.ctypes.data_as(C.c_void_p), *gptr)
    if rc:
        raise TpcfError(rc, tape.ctx.last_error())


# ------------------------------------------------------------ bootstrap ----
def block_sums(res: "Shear2pcfResults", patch_A, w_A, n_patch_A: int) -> dict:
    """Per lens patch ("block") sums of a correlation with per-patch-pair
    outputs: G_t, G_x, K [n_block][n_bin] (the patch-pair sums over source
    patches, which equal the block sums of the marks, src/tpcf.f90:1130-1137)
    and the block weights (sum of lens w).  Empty blocks (weight 0) are
    dropped, and a block's G_t, G_x are set to 0 in the bins where its K is 0,
    as src/tpcf.f90:1054-1069 does."""
    gt = np.asarray(res.pp_gt).sum(axis=0)
    gx = np.asarray(res.pp_gx).sum(axis=0)
    k = np.asarray(res.pp_k).sum(axis=0)
    w = np.bincount(np.asarray(patch_A), weights=np.asarray(w_A, dtype=np.float64), minlength=n_patch_A)
    keep = w > 0
    gt, gx, k = gt[keep], gx[keep], k[keep]
    gt = np.where(k == 0, 0.0, gt)
    gx = np.where(k == 0, 0.0, gx)
    return dict(g_t=gt, g_x=gx, k=k, weight=w[keep])


def rect_blocks(x, y, n_x_blocks: int, n_y_blocks: int, *, offset: int = 0, device_inputs: bool = False,
                labels=None, n: Optional[int] = None, precision: Optional[str] = None,
                ctx: Optional[Context] = None):
    """Rectangular bootstrap block labels of one field on the GPU
    (tpcf_rect_blocks_*): the reference's default block scheme,
    src/tpcf.f90:1017-1049.  0-based labels in [offset, offset + n_x_blocks
    n_y_blocks).  Host arrays in, an int32 array out; with ``device_inputs``
    x, y and ``labels`` are device addresses (``n`` objects, ``precision``
    "f32" or "f64")."""
    ctx = ctx or default_context()
    if device_inputs:
        if labels is None or n is None or precision is None:
            raise ValueError("device_inputs needs labels, n and precision")
        fn = N.lib().tpcf_rect_blocks_f32 if precision == "f32" else N.lib().tpcf_rect_blocks_f64
        rc = fn(ctx.handle, int(n), x, y, int(n_x_blocks), int(n_y_blocks), int(offset), N.FLAG_DEVICE_INPUTS, labels)
        if rc:
            raise TpcfError(rc, ctx.last_error())
        return labels
    prec = precision or _precision_of(x)
    dt = np.float32 if prec == "f32" else np.float64
    xs, ys = _host(x, dt), _host(y, dt)
    out = np.zeros(len(xs), dtype=np.int32)
    fn = N.lib().tpcf_rect_blocks_f32 if prec == "f32" else N.lib().tpcf_rect_blocks_f64
    rc = fn(ctx.handle, len(xs), xs.ctypes.data, ys.ctypes.data, int(n_x_blocks), int(n_y_blocks), int(offset), 0,
            out.ctypes.data)
    if rc:
        raise TpcfError(rc, ctx.last_error())
    return out


def bootstrap(blocks: dict, n_boot: int, *, multiplicity=None, seed: int = 0, supersampling: float = 1.0,
              covariance: bool = True, ctx: Optional[Context] = None) -> dict:
    """Block bootstrap on the GPU (tpcf_bootstrap): the resampled g_t, g_x
    [n_boot][n_bin], weights [n_boot], the multiplicities [n_boot][n_block] and
    the covariances of g_t and g_x [n_bin][n_bin] (src/tpcf.f90:1165-1191,
    src/utils.f90:745-764).  ``blocks``: g_t, g_x, k [n_block][n_bin] and
    optional weight [n_block] (see block_sums)."""
    ctx = ctx or default_context()
    gt = np.ascontiguousarray(blocks["g_t"], dtype=np.float64)
    gx = np.ascontiguousarray(blocks["g_x"], dtype=np.float64)
    k = np.ascontiguousarray(blocks["k"], dtype=np.float64)
    w = blocks.get("weight")
    w = None if w is None else np.ascontiguousarray(w, dtype=np.float64)
    P, nb = gt.shape
    b = N.Blocks()
    b.n_block, b.n_bin = P, nb
    b.g_t, b.g_x, b.k = gt.ctypes.data, gx.ctypes.data, k.ctypes.data
    b.weight = w.ctypes.data if w is not None else None
    m_in = None
    if multiplicity is not None:
        m_in = np.ascontiguousarray(multiplicity, dtype=np.int32)
        if m_in.shape != (n_boot, P):
            raise ValueError("multiplicity must be [n_boot][n_block]")
    out = dict(g_t=np.zeros((n_boot, nb)), g_x=np.zeros((n_boot, nb)), weights=np.zeros(n_boot),
               multiplicity=np.zeros((n_boot, P), dtype=np.int32))
    if covariance:
        out["cov_t"] = np.zeros((nb, nb))
        out["cov_x"] = np.zeros((nb, nb))
    ptr = lambda a: None if a is None else a.ctypes.data   # noqa: E731
    rc = N.lib().tpcf_bootstrap(ctx.handle, C.byref(b), int(n_boot), ptr(m_in), int(seed) & ((1 << 64) - 1),
                                float(supersampling), 0, ptr(out["g_t"]), ptr(out["g_x"]), ptr(out["weights"]),
                                ptr(out.get("cov_t")), ptr(out.get("cov_x")), ptr(out["multiplicity"]))
    if rc:
        raise TpcfError(rc, ctx.last_error())
    return out


# --------------------------------------------------------------- k-means ----
def kmeans(x, y, w, k: int, *, max_iterations: int = 1000, reassignment_tol: float = 0.0, centroids=None,
           spherical: bool = False, weight_equalizer_scaling: float = 0.0, equal_size: bool = False,
           weight_tol: float = 0.0, precision: Optional[str] = None, ctx: Optional[Context] = None) -> dict:
    """k-means patch labels on the GPU (tpcf_kmeans_*): partition_data
    (src/kmeans.f90:12-263).  Returns 0-based labels, centroids [k][2],
    cluster weights, the rounds run and the points moved in the last round.
    ``centroids`` (k x 2) replaces the fixed starting grid.
    ``weight_equalizer_scaling`` > 0 scales the distances by the clusters'
    weights, moved after every step of 256 consecutive points (the reference
    moves them after every point); ``equal_size`` adds the post-processing
    that hands points to their runner-up clusters (``weight_tol``)."""
    ctx = ctx or default_context()
    prec = precision or _precision_of(x)
    dt = np.float32 if prec == "f32" else np.float64
    xs, ys, ws = (np.ascontiguousarray(v, dtype=dt) for v in (x, y, w))
    n = len(xs)
    o = N.KmeansOpts()
    o.k, o.max_iterations, o.reassignment_tol = int(k), int(max_iterations), float(reassignment_tol)
    o.init = 0 if centroids is None else 1
    o.weight_equalizer_scaling = float(weight_equalizer_scaling)
    o.equal_size = 1 if equal_size else 0
    o.weight_tol = float(weight_tol)
    cen = np.zeros((k, 2)) if centroids is None else np.array(centroids, dtype=np.float64).reshape(k, 2)
    lab = np.zeros(n, dtype=np.int32)
    cw = np.zeros(k)
    rounds, moves = C.c_int32(0), C.c_int64(0)
    fn = N.lib().tpcf_kmeans_f32 if prec == "f32" else N.lib().tpcf_kmeans_f64
    rc = fn(ctx.handle, n, xs.ctypes.data, ys.ctypes.data, ws.ctypes.data, C.byref(o),
            N.FLAG_SPHERICAL if spherical else 0, lab.ctypes.data, cen.ctypes.data, cw.ctypes.data,
            C.byref(rounds), C.byref(moves))
    if rc:
        raise TpcfError(rc, ctx.last_error())
    return dict(labels=lab, centroids=cen, cluster_weights=cw, iterations=rounds.value, last_moves=moves.value)


# ---------------------------------------------------------------- batch ----
class Batch:
    """Many catalogues run as one HIP graph (tpcf_batch_*).

    ``catalogs``: list of dicts with ``kind`` (SHEAR / KAPPA), ``lens`` and
    ``src`` (device addresses of arrays of the batch precision: lens x, y,
    (z,) val, w; source x, y, (z,) e1/e2 or kappa_E/kappa_B, w, (m)), sizes
    ``n_A``, ``n_B`` and optional ``patch_A``, ``patch_B`` (device int32
    addresses) with ``n_patch_A``, ``n_patch_B``.  The arrays must stay valid
    while the batch exists; ``run()`` re-reads their values each time (the
    positions, w == 0 lenses and z < 0 masks are fixed at creation).  With
    ``marks`` (True, or a collection of catalogue indices) each result (of
    those catalogues) also holds the per-lens marks as torch tensors on the
    GPU, (n_bin, n_A) views of [n_A][n_bin] FP64 arrays, rewritten by every
    run; with ``spherical`` x, y are longitude and latitude and the bins
    angles, all in radians."""

    def __init__(self, catalogs: list, bins: Binning, *, precision: str = "f32", autotune: bool = False,
                 graph: bool = True, patches: bool = False, ctx: Optional[Context] = None, marks: bool = False,
                 spherical: bool = False):
        self.ctx = ctx or default_context()
        self.bins = bins
        self._h = C.c_void_p()
        n = len(catalogs)
        self._cats = (N.Catalog * max(n, 1))()
        self._outs = []
        nb = bins.n_bin
        for i, cdesc in enumerate(catalogs):
            kind = cdesc["kind"]
            L, S = cdesc["lens"], cdesc["src"]
            s1, s2 = ("e1", "e2") if kind == SHEAR else ("kappa_E", "kappa_B")
            c = self._cats[i]
            c.kind = kind
            c.n_A, c.n_B = int(cdesc["n_A"]), int(cdesc["n_B"])
            c.x_A, c.y_A, c.val_A, c.w_A = L["x"], L["y"], L["val"], L["w"]
            c.z_A = L.get("z") if kind == SHEAR else None
            c.x_B, c.y_B, c.s1_B, c.s2_B, c.w_B = S["x"], S["y"], S[s1], S[s2], S["w"]
            c.z_B = S.get("z") if kind == SHEAR else None
            c.m_B = S.get("m") if kind == SHEAR else None
            c.patch_A, c.patch_B = cdesc.get("patch_A"), cdesc.get("patch_B")
            c.n_patch_A, c.n_patch_B = int(cdesc.get("n_patch_A", 1)), int(cdesc.get("n_patch_B", 1))
            res = {k: np.zeros(nb, dtype=np.float64) for k in ("theta", "mean_theta", "g_t", "g_x", "weights",
                                                                "sum_theta", "sum_gt", "sum_gx", "sum_w", "sum_k")}
            res["n_pairs"] = np.zeros(nb, dtype=np.int64)
            o = N.Out()
            o.real_bytes = 8
            for k, v in res.items():
                setattr(o, k, v.ctypes.data_as(C.c_void_p))
            if patches:
                pa = c.n_patch_A if c.patch_A else 1
                pb = c.n_patch_B if c.patch_B else 1
                res["pp_n"] = np.zeros((pb, pa, nb), dtype=np.int64)
                o.pp_n = res["pp_n"].ctypes.data_as(C.c_void_p)
                for k in ("pp_theta", "pp_gt", "pp_gx", "pp_w", "pp_k"):
                    res[k] = np.zeros((pb, pa, nb), dtype=np.float64)
                    setattr(o, k, res[k].ctypes.data_as(C.c_void_p))
            if marks is True or (marks and not isinstance(marks, bool) and i in marks):
                import torch
                dev = torch.device("cuda", self.ctx.device) if self.ctx.device >= 0 else torch.device("cuda")
                for k in ("xi_E_marks", "xi_B_marks", "K_marks", "w_marks"):
                    res[k] = torch.zeros((c.n_A, nb), dtype=torch.float64, device=dev)
                    setattr(o, k, C.c_void_p(res[k].data_ptr()))
                res["marks_weight"] = torch.zeros(c.n_A, dtype=torch.float64, device=dev)
                o.marks_weight = C.c_void_p(res["marks_weight"].data_ptr())
                torch.cuda.synchronize(dev)
            self._outs.append((o, res))
            c.out = C.pointer(o)
        flags = N.FLAG_DEVICE_INPUTS | (N.FLAG_BATCH_AUTOTUNE if autotune else 0) | \
            (0 if graph else N.FLAG_BATCH_NO_GRAPH) | (N.FLAG_SPHERICAL if spherical else 0)
        fn = N.lib().tpcf_batch_create_f32 if precision == "f32" else N.lib().tpcf_batch_create_f64
        rc = fn(self.ctx.handle, n, self._cats, C.byref(bins.c), flags, C.byref(self._h))
        if rc:
            raise TpcfError(rc, self.ctx.last_error())

    def run(self) -> list:
        """One run of every catalogue: a list of result dicts (per-bin
        outputs, raw sums, per-patch-pair sums when asked for)."""
        if not self._h:
            raise TpcfError(N.ERR_ARG, "the batch is closed")
        if not self.ctx.handle:
            raise TpcfError(N.ERR_ARG, "the batch's context is closed")
        rc = N.lib().tpcf_batch_run(self._h)
        if rc:
            raise TpcfError(rc, self.ctx.last_error())
        out = []
        for _, res in self._outs:
            r = {}
            for k, v in res.items():
                if k.endswith("marks"):
                    r[k] = v.clone().T         # (n_bin, n_A) view of a copy of the [n_A][n_bin] array
                elif k == "marks_weight":
                    r[k] = v.clone()
                else:
                    r[k] = v.copy()
            out.append(r)
        return out

    @property
    def stats(self) -> dict:
        st = N.BatchStats()
        N.lib().tpcf_batch_get_stats(self._h, C.byref(st))
        return st.as_dict()

    def close(self) -> None:
        if self._h:
            N.lib().tpcf_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
