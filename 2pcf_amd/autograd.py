"""torch.autograd.Function wrappers of the correlation and its adjoint.

    g_t, g_x = shear_scalar(lens, src, bins)      # lens: x, y, z, val, w; src: x, y, z, e1, e2, w, m
    xi_E, xi_B = kappa_kappa(lens, src, bins)     # lens: x, y, val, w; src: x, y, kappa_E, kappa_B, w

Arguments are CUDA (HIP) tensors of one float dtype (float32: the FP32 path,
float64: the FP64 path).  The outputs are float64 tensors [n_bin] on the same
device.  The forward records the library's lens-side adjoint pass
(tpcf_adjoint_record_*), whose per-lens sums also give the outputs; backward
replays it (tpcf_adjoint_replay: the source-side pass only), so a forward and
a backward cost two pair passes.  Backward returns gradients for val and w of
the lenses and for the value components, w (and m) of the sources; positions
and z get none.  The same tensor may be passed on both sides
(auto-correlation): autograd adds the two contributions.  With
``spherical=True`` x and y are longitude and latitude and the bin edges angles,
all in radians (the shear angle is the one at the lens, as in the forward).
"""
from __future__ import annotations

import torch

from . import tpcf as T

_LENS = {T.SHEAR: ("x", "y", "z", "val", "w"), T.KAPPA: ("x", "y", "val", "w")}
_SRC = {T.SHEAR: ("x", "y", "z", "e1", "e2", "w", "m"), T.KAPPA: ("x", "y", "kappa_E", "kappa_B", "w")}
_DIFF_LENS = ("val", "w")
_DIFF_SRC = {T.SHEAR: ("e1", "e2", "w", "m"), T.KAPPA: ("kappa_E", "kappa_B", "w")}


def _prec(t: torch.Tensor) -> str:
    if t.dtype == torch.float32:
        return "f32"
    if t.dtype == torch.float64:
        return "f64"
    raise TypeError("inputs must be float32 or float64 tensors")


class _Correlation(torch.autograd.Function):
    @staticmethod
    def forward(ctx, kind, bins, tctx, spherical, *tensors):
        lk, sk = _LENS[kind], _SRC[kind]
        lens = dict(zip(lk, tensors[:len(lk)]))
        src = dict(zip(sk, tensors[len(lk):]))
        for t in tensors:
            if not t.is_cuda or not t.is_contiguous():
                raise ValueError("inputs must be contiguous device tensors")
        prec = _prec(lens["x"])
        torch.cuda.current_stream().synchronize()
        g1, g2, tape = T.adjoint_record(kind, {k: v.data_ptr() for k, v in lens.items()},
                                        {k: v.data_ptr() for k, v in src.items()}, bins, precision=prec, ctx=tctx,
                                        n_A=lens["x"].numel(), n_B=src["x"].numel(), spherical=spherical)
        ctx.kind, ctx.bins, ctx.tape = kind, bins, tape
        # keeps the inputs alive (the tape holds their addresses) and detects
        # their modification in place before backward
        ctx.save_for_backward(*tensors)
        dev = lens["x"].device
        return torch.as_tensor(g1, device=dev), torch.as_tensor(g2, device=dev)

    @staticmethod
    def backward(ctx, gbar1, gbar2):
        kind = ctx.kind
        tensors = ctx.saved_tensors
        lk, sk = _LENS[kind], _SRC[kind]
        lens = dict(zip(lk, tensors[:len(lk)]))
        src = dict(zip(sk, tensors[len(lk):]))
        needs = ctx.needs_input_grad[4:]
        need = {("lens", k): needs[i] for i, k in enumerate(lk)}
        need.update({("src", k): needs[len(lk) + i] for i, k in enumerate(sk)})
        dev = lens["x"].device
        grads, bufs = {"lens": {}, "src": {}}, {}
        for side, keys, n in (("lens", _DIFF_LENS, lens["x"].numel()), ("src", _DIFF_SRC[kind], src["x"].numel())):
            for k in keys:
                if need.get((side, k)):
                    b = torch.empty(n, dtype=torch.float64, device=dev)
                    bufs[(side, k)] = b
                    grads[side][k] = b.data_ptr()
        g1 = (gbar1 if gbar1 is not None else torch.zeros(ctx.bins.n_bin)).detach().double().cpu().numpy()
        g2 = (gbar2 if gbar2 is not None else torch.zeros(ctx.bins.n_bin)).detach().double().cpu().numpy()
        if bufs:
            torch.cuda.current_stream().synchronize()
            T.adjoint_replay(ctx.tape, g1, g2, grads)
        out = []
        for side, keys, tensors_ in (("lens", lk, lens), ("src", sk, src)):
            for k in keys:
                b = bufs.get((side, k))
                out.append(b.to(tensors_[k].dtype) if b is not None else None)
        return (None, None, None, None, *out)


def shear_scalar(lens: dict, src: dict, bins, ctx=None, spherical: bool = False):
    """(g_t, g_x) per bin of tangential_shear_tpcf (src/tpcf.f90:190-378), differentiable."""
    ctx = ctx or T.default_context()
    args = [lens[k] for k in _LENS[T.SHEAR]] + [src[k] for k in _SRC[T.SHEAR]]
    return _Correlation.apply(T.SHEAR, bins, ctx, bool(spherical), *args)


def kappa_kappa(lens: dict, src: dict, bins, ctx=None, spherical: bool = False):
    """(xi_E, xi_B) per bin of scalar_kappa_tpcf (src/tpcf.f90:380-545), differentiable."""
    ctx = ctx or T.default_context()
    args = [lens[k] for k in _LENS[T.KAPPA]] + [src[k] for k in _SRC[T.KAPPA]]
    return _Correlation.apply(T.KAPPA, bins, ctx, bool(spherical), *args)


def chi2(d: torch.Tensor, model: torch.Tensor, sigma: torch.Tensor) -> torch.Tensor:
    """(d - m)^T C^-1 (d - m) with diagonal C = diag(sigma^2)."""
    r = (d - model) / sigma
    return (r * r).sum()
