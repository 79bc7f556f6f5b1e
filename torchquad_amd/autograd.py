"""Differentiable parameter sweeps with a hand-written backward kernel.

``gaussian_peak_integrals(a, u, domains, N, rule)`` integrates
f_b(x) = exp(-sum_k a_bk^2 (x_k - u_bk)^2) over the box domains[b] for every
row b with one forward kernel, and its backward is one kernel that returns
d/da, d/du and d/d(domain bounds) (see tq_sweep_gaussian_backward).  It
replaces a Python loop of integrate() calls differentiated by torch autograd
through the materialised grid (the reference's way, gaussian.py:44-162,
tests/gradient_test.py).
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


class _GaussianPeakSweep(torch.autograd.Function):
    @staticmethod
    def forward(ctx, a, u, domains, rule, n):
        B, dim = a.shape
        g, gl = _desc(rule, dim, n, a.dtype, a.device)
        nbytes = lib().tq_sweep_workspace_size(ctypes.byref(g), B)
        ws = workspace.get(max(nbytes, 8), a.device, tag="sweep")
        out = torch.empty(B, dtype=a.dtype, device=a.device)
        call("tq_sweep_gaussian_forward", ctypes.byref(g), B, ptr(a), ptr(u), ptr(domains), ptr(ws), ws.numel(),
             ptr(out), None, stream_ptr(a.device))
        ctx.save_for_backward(a, u, domains)
        ctx.meta = (rule, n, gl)
        return out

    @staticmethod
    def backward(ctx, grad_out):
        a, u, domains = ctx.saved_tensors
        rule, n, gl = ctx.meta
        B, dim = a.shape
        g = _ops.grid_desc(rule, dim, n, a.dtype, gl)
        nbytes = lib().tq_sweep_workspace_size(ctypes.byref(g), B)
        ws = workspace.get(max(nbytes, 8), a.device, tag="sweep")
        grad_out = grad_out.contiguous().to(a.dtype)
        ga = torch.empty_like(a) if ctx.needs_input_grad[0] else None
        gu = torch.empty_like(u) if ctx.needs_input_grad[1] else None
        gd = torch.empty_like(domains) if ctx.needs_input_grad[2] else None
        call("tq_sweep_gaussian_backward", ctypes.byref(g), B, ptr(a), ptr(u), ptr(domains), ptr(grad_out), ptr(ws),
             ws.numel(), ptr(ga), ptr(gu), ptr(gd), stream_ptr(a.device))
        return ga, gu, gd, None, None


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
    require_device(a)
    if a.dim() != 2 or u.shape != a.shape or domains.shape != (*a.shape, 2):
        raise ValueError("expected a, u of shape [B, dim] and domains of shape [B, dim, 2]")
    if a.shape[1] > 8:
        raise ValueError("gaussian_peak_integrals supports dim <= 8")
    code = RULE_NAMES[rule.lower()] if isinstance(rule, str) else int(rule)
    if n is None:
        if N is None:
            raise ValueError("give N (total points) or n (points per axis)")
        n = n_per_axis(code, a.shape[1], int(N))
    dtype = a.dtype
    u = u.to(dtype).contiguous()
    domains = domains.to(dtype).contiguous()
    return _GaussianPeakSweep.apply(a.contiguous(), u, domains, code, int(n))
