"""Differentiable parameter sweeps with hand-written backward kernels.

``gaussian_peak_integrals(a, u, domains, N, rule)`` integrates
f_b(x) = exp(-sum_k a_bk^2 (x_k - u_bk)^2) over the box domains[b] for every
row b with one forward kernel, and its backward is one kernel that returns
d/da, d/du and d/d(domain bounds) (tq_sweep_backward);
``oscillatory_integrals(c, w, ...)`` and ``product_peak_integrals(a, u, ...)``
do the same for the other two Genz families of the batched workload.  They
replace a Python loop of integrate() calls differentiated by torch autograd
through the materialised grid (the reference's way, gaussian.py:44-162,
utils.py:24-58, tests/gradient_test.py:162-259).
"""
import ctypes

import torch

from . import _lib, _ops
from ._lib import call, lib, ptr, require_device, stream_ptr, workspace
from .batch import RULE_NAMES, n_per_axis
from .gl_nodes import gauss_legendre

_gl = {}


def _gl_tables(n, device):
    key = (int(n), str(device))
    if key not in _gl:
        _gl[key] = gauss_legendre(int(n), device=device)
    return _gl[key]


def _desc(rule, dim, n, dtype, device):
    gl = _gl_tables(n, device) if rule == _lib.RULE_GAUSS_LEGENDRE else None
    return _ops.grid_desc(rule, dim, n, dtype, gl), gl


class _Sweep(torch.autograd.Function):
    """Forward and hand-written backward of one family's parameter sweep."""

    @staticmethod
    def forward(ctx, a, u, domains, family, rule, n):
        B, dim = a.shape
        g, gl = _desc(rule, dim, n, a.dtype, a.device)
        nbytes = lib().tq_sweep_workspace_size(ctypes.byref(g), B)
        ws = workspace.get(max(nbytes, 8), a.device, tag="sweep")
        out = torch.empty(B, dtype=a.dtype, device=a.device)
        call("tq_sweep_forward", ctypes.byref(g), family, B, ptr(a), ptr(u), ptr(domains), ptr(ws), ws.numel(),
             ptr(out), None, stream_ptr(a.device))
        ctx.save_for_backward(a, u, domains)
        ctx.meta = (family, rule, n, gl)
        return out

    @staticmethod
    def backward(ctx, grad_out):
        a, u, domains = ctx.saved_tensors
        family, rule, n, gl = ctx.meta
        B, dim = a.shape
        g = _ops.grid_desc(rule, dim, n, a.dtype, gl)
        nbytes = lib().tq_sweep_workspace_size(ctypes.byref(g), B)
        ws = workspace.get(max(nbytes, 8), a.device, tag="sweep")
        grad_out = grad_out.contiguous().to(a.dtype)
        ga = torch.empty_like(a) if ctx.needs_input_grad[0] else None
        gu = torch.empty_like(u) if ctx.needs_input_grad[1] else None
        gd = torch.empty_like(domains) if ctx.needs_input_grad[2] else None
        call("tq_sweep_backward", ctypes.byref(g), family, B, ptr(a), ptr(u), ptr(domains), ptr(grad_out), ptr(ws),
             ws.numel(), ptr(ga), ptr(gu), ptr(gd), stream_ptr(a.device))
        return ga, gu, gd, None, None, None


def _sweep(family, a, u, domains, N, n, rule, name, u_per_axis=True):
    require_device(a)
    if a.dim() != 2 or domains.shape != (*a.shape, 2):
        raise ValueError(f"{name}: expected parameters of shape [B, dim] and domains of shape [B, dim, 2]")
    if (u_per_axis and u.shape != a.shape) or (not u_per_axis and u.shape != a.shape[:1]):
        raise ValueError(f"{name}: second parameter has shape {tuple(u.shape)}")
    if a.shape[1] > 8:
        raise ValueError(f"{name} supports dim <= 8")
    code = RULE_NAMES[rule.lower()] if isinstance(rule, str) else int(rule)
    if n is None:
        if N is None:
            raise ValueError("give N (total points) or n (points per axis)")
        n = n_per_axis(code, a.shape[1], int(N))
    dtype = a.dtype
    u = u.to(dtype).contiguous()
    domains = domains.to(dtype).contiguous()
    return _Sweep.apply(a.contiguous(), u, domains, family, code, int(n))


def gaussian_peak_integrals(a, u, domains, N=None, n=None, rule="gauss_legendre"):
    """Integrals of exp(-sum a^2 (x-u)^2) over domains[b] for each row b.

    Args:
        a, u: [B, dim] tensors (float32 or float64, on the GPU).
        domains: [B, dim, 2] tensor of [lo, hi] per axis.
        N: total points per integral (rounded per axis like the integrators),
            or n: points per axis.
        rule: "gauss_legendre", "trapezoid", "simpson" or "boole".

    Returns:
        [B] tensor; differentiable w.r.t. a, u and domains.
    """
    return _sweep(_lib.FAM_GAUSSIAN_PEAK, a, u, domains, N, n, rule, "gaussian_peak_integrals")


def oscillatory_integrals(c, w, domains, N=None, n=None, rule="gauss_legendre"):
    """Integrals of cos(2 pi w + sum_k c_k x_k) over domains[b] for each row b
    (Genz's oscillatory family, as ``integrands.Oscillatory``).

    Args:
        c: [B, dim]; w: [B]; domains: [B, dim, 2] (float32 or float64, GPU).
        N / n, rule: as for :func:`gaussian_peak_integrals`.

    Returns:
        [B] tensor; differentiable w.r.t. c, w and domains.
    """
    return _sweep(_lib.FAM_OSCILLATORY, c, w, domains, N, n, rule, "oscillatory_integrals", u_per_axis=False)


def product_peak_integrals(a, u, domains, N=None, n=None, rule="gauss_legendre"):
    """Integrals of prod_k 1/(a_k^-2 + (x_k - u_k)^2) over domains[b] for each
    row b (Genz's product-peak family, as ``integrands.ProductPeak``).

    Args and returns as :func:`gaussian_peak_integrals`; differentiable w.r.t.
    a, u and domains.
    """
    return _sweep(_lib.FAM_PRODUCT_PEAK, a, u, domains, N, n, rule, "product_peak_integrals")
