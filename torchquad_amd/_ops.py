"""Tensor-level wrappers of the C ABI: grid, Monte Carlo and RNG kernels.

Every function here launches HIP kernels on torch's current stream of the
domain's device and returns device tensors without synchronising.
"""
import ctypes

import torch

from . import _lib
from ._lib import call, lib, ptr, require_device, stream_ptr, workspace

RULES = {
    "trapezoid": _lib.RULE_TRAPEZOID,
    "simpson": _lib.RULE_SIMPSON,
    "boole": _lib.RULE_BOOLE,
    "gauss_legendre": _lib.RULE_GAUSS_LEGENDRE,
}


def grid_desc(rule, dim, n, dtype, gl=None, lane_axes=0, items_per_wave=0):
    g = _lib.GridDesc()
    g.dtype = _lib.dtype_code(dtype)
    g.rule = int(rule)
    g.dim = int(dim)
    g.n = int(n)
    g.lane_axes = int(lane_axes)
    g.items_per_wave = int(items_per_wave)
    if gl is not None:
        g.gl_nodes = gl[0].data_ptr()
        g.gl_weights = gl[1].data_ptr()
    return g


def _fused_grid_desc(rule, integrand, dim, n, dtype, gl, lane_axes, items_per_wave):
    desc = integrand.descriptor(dim)
    if items_per_wave == 0 and lane_axes == 0:
        from .autotune import grid_ipw

        items_per_wave = grid_ipw(rule, desc.family, dim, n, dtype)
    return grid_desc(rule, dim, n, dtype, gl, lane_axes, items_per_wave), desc


def grid_workspace(rule, integrand, dim, n, domain, gl=None, lane_axes=0, items_per_wave=0):
    """A new workspace tensor sized for grid_integrate with these arguments
    (for callers that must own theirs, e.g. captured graphs)."""
    g, _ = _fused_grid_desc(rule, integrand, dim, n, domain.dtype, gl, lane_axes, items_per_wave)
    return torch.zeros(max(_grid_ws_bytes(g), 8), dtype=torch.uint8, device=domain.device)


_ws_sizes = {}


def _grid_ws_bytes(g):
    """tq_grid_workspace_size, remembered per geometry (it depends on the
    descriptor's scalar fields only)."""
    key = (g.dtype, g.rule, g.dim, g.n, g.lane_axes, g.items_per_wave)
    nbytes = _ws_sizes.get(key)
    if nbytes is None:
        nbytes = int(lib().tq_grid_workspace_size(ctypes.byref(g)))
        _ws_sizes[key] = nbytes
    return nbytes


def grid_integrate(rule, integrand, dim, n, domain, gl=None, lane_axes=0, items_per_wave=0, out_f64=None, ws=None):
    """Fused tensor-product rule for a built-in integrand; 0-dim result tensor.

    ws: the caller's workspace (uint8 tensor); default the per-stream shared one."""
    require_device(domain)
    g, desc = _fused_grid_desc(rule, integrand, dim, n, domain.dtype, gl, lane_axes, items_per_wave)
    nbytes = _grid_ws_bytes(g)
    stream = stream_ptr(domain.device)
    if ws is None:
        ws = workspace.get(max(nbytes, 8), domain.device, stream=stream)
    elif ws.numel() < nbytes:
        raise ValueError(f"workspace of {ws.numel()} bytes, {nbytes} needed")
    out = torch.empty((), dtype=domain.dtype, device=domain.device)
    call(
        "tq_grid_integrate",
        ctypes.byref(g),
        ctypes.byref(desc),
        ptr(domain),
        ptr(ws),
        ws.numel(),
        ptr(out),
        ptr(out_f64),
        stream,
    )
    return out


def grid_points(rule, dim, n, domain, gl=None):
    """[n^dim, dim] grid points (ij order) written by the HIP point writer."""
    require_device(domain)
    g = grid_desc(rule, dim, n, domain.dtype, gl)
    pts = torch.empty((n**dim, dim), dtype=domain.dtype, device=domain.device)
    call("tq_grid_points", ctypes.byref(g), ptr(domain), ptr(pts), stream_ptr(domain.device))
    return pts


def grid_h(rule, dim, n, domain, gl=None):
    require_device(domain)
    g = grid_desc(rule, dim, n, domain.dtype, gl)
    h = torch.empty((dim,), dtype=domain.dtype, device=domain.device)
    call("tq_grid_h", ctypes.byref(g), ptr(domain), ptr(h), stream_ptr(domain.device))
    return h


def grid_reduce(rule, dim, n, domain, values, hs=None, unit_weights=False, gl=None):
    """Weighted tensor-product sum of values[n^dim, *S] -> [*S]."""
    require_device(domain)
    values = values.contiguous()
    if values.dtype != domain.dtype:
        values = values.to(domain.dtype)
    S = 1
    for s in values.shape[1:]:
        S *= int(s)
    g = grid_desc(rule, dim, n, domain.dtype, gl)
    nbytes = lib().tq_grid_reduce_workspace_size(ctypes.byref(g), S)
    ws = workspace.get(max(nbytes, 8), domain.device, tag="reduce")
    out = torch.empty(tuple(values.shape[1:]), dtype=domain.dtype, device=domain.device)
    if hs is not None:
        hs = hs.detach().to(domain.dtype).contiguous()
    call(
        "tq_grid_reduce",
        ctypes.byref(g),
        ptr(domain),
        ptr(hs),
        1 if unit_weights else 0,
        ptr(values),
        S,
        ptr(ws),
        ws.numel(),
        ptr(out),
        stream_ptr(domain.device),
    )
    return out


def grid_reduce_backward(rule, dim, n, domain, grad_out, npoints, hs=None, unit_weights=False, gl=None):
    require_device(domain)
    grad_out = grad_out.contiguous().to(domain.dtype)
    S = grad_out.numel()
    g = grid_desc(rule, dim, n, domain.dtype, gl)
    grad_f = torch.empty((npoints, *grad_out.shape), dtype=domain.dtype, device=domain.device)
    if hs is not None:
        hs = hs.detach().to(domain.dtype).contiguous()
    call(
        "tq_grid_reduce_backward",
        ctypes.byref(g),
        ptr(domain),
        ptr(hs),
        1 if unit_weights else 0,
        ptr(grad_out),
        S,
        ptr(grad_f),
        stream_ptr(domain.device),
    )
    return grad_f


def grid_points_backward(rule, dim, n, domain, grad_points, gl=None):
    require_device(domain)
    g = grid_desc(rule, dim, n, domain.dtype, gl)
    grad_points = grad_points.contiguous().to(domain.dtype)
    nbytes = lib().tq_grid_points_backward_workspace_size(ctypes.byref(g))
    ws = workspace.get(max(nbytes, 8), domain.device, tag="pbwd")
    gd = torch.empty((dim, 2), dtype=domain.dtype, device=domain.device)
    call(
        "tq_grid_points_backward",
        ctypes.byref(g),
        ptr(domain),
        ptr(grad_points),
        ptr(ws),
        ws.numel(),
        ptr(gd),
        stream_ptr(domain.device),
    )
    return gd


# ---------------------------------------------------------------------------
# Monte Carlo and RNG
# ---------------------------------------------------------------------------
def mc_desc(dim, N, dtype, seed, offset, samples_per_thread=0):
    m = _lib.McDesc()
    m.dtype = _lib.dtype_code(dtype)
    m.dim = int(dim)
    m.N = int(N)
    m.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    m.offset = int(offset) & 0xFFFFFFFFFFFFFFFF
    m.samples_per_thread = int(samples_per_thread)
    return m


def mc_workspace(dim, N, domain):
    """A new workspace tensor for mc_integrate (for callers that own theirs)."""
    m = mc_desc(dim, N, domain.dtype, 0, 0)
    return torch.empty(max(int(lib().tq_mc_workspace_size(ctypes.byref(m))), 16), dtype=torch.uint8,
                       device=domain.device)


def mc_integrate(integrand, dim, N, domain, seed, offset=0, stats=None, samples_per_thread=0, desc=None,
                 offset_dev=None, ws=None, out=None):
    """Fused Monte Carlo; 0-dim result.  desc: a C descriptor instead of
    integrand's; offset_dev: int64 device tensor holding the first Philox
    counter (then `offset` is ignored); ws / out: caller-owned buffers."""
    require_device(domain)
    m = mc_desc(dim, N, domain.dtype, seed, offset, samples_per_thread)
    if desc is None:
        desc = integrand.descriptor(dim)
    nbytes = lib().tq_mc_workspace_size(ctypes.byref(m))
    if ws is None:
        ws = workspace.get(max(nbytes, 16), domain.device, tag="mc")
    elif ws.numel() < nbytes:
        raise ValueError(f"workspace of {ws.numel()} bytes, {nbytes} needed")
    if out is None:
        out = torch.empty((), dtype=domain.dtype, device=domain.device)
    args = (ctypes.byref(m), ctypes.byref(desc), ptr(domain))
    tail = (ptr(ws), ws.numel(), ptr(out), ptr(stats), stream_ptr(domain.device))
    if offset_dev is not None:
        call("tq_mc_integrate_dev_offset", *args, ptr(offset_dev), *tail)
    else:
        call("tq_mc_integrate", *args, *tail)
    return out


def mc_points(dim, N, domain, seed, offset=0):
    require_device(domain)
    m = mc_desc(dim, N, domain.dtype, seed, offset)
    pts = torch.empty((int(N), dim), dtype=domain.dtype, device=domain.device)
    call("tq_mc_points", ctypes.byref(m), ptr(domain), ptr(pts), stream_ptr(domain.device))
    return pts


def mc_reduce(values, domain):
    """volume * mean over the first axis of values[N, *S]."""
    require_device(domain)
    values = values.contiguous()
    if values.dtype != domain.dtype:
        values = values.to(domain.dtype)
    N = int(values.shape[0])
    S = 1
    for s in values.shape[1:]:
        S *= int(s)
    nbytes = lib().tq_mc_reduce_workspace_size(N, S)
    ws = workspace.get(max(nbytes, 8), domain.device, tag="reduce")
    out = torch.empty(tuple(values.shape[1:]), dtype=domain.dtype, device=domain.device)
    call(
        "tq_mc_reduce",
        _lib.dtype_code(domain.dtype),
        ptr(values),
        N,
        S,
        ptr(domain),
        int(domain.shape[0]),
        ptr(ws),
        ws.numel(),
        ptr(out),
        stream_ptr(domain.device),
    )
    return out


def uniform(seed, offset, rows, cols, dtype, device):
    """rows x cols uniforms in [0,1) from the library's Philox stream."""
    out = torch.empty((int(rows), int(cols)), dtype=dtype, device=device)
    if rows == 0:
        return out
    if device.type != "cuda" or not torch.cuda.is_available():
        raise _lib.NativeUnavailableError("random numbers are generated on a ROCm GPU only")
    call(
        "tq_uniform",
        int(seed) & 0xFFFFFFFFFFFFFFFF,
        int(offset) & 0xFFFFFFFFFFFFFFFF,
        _lib.dtype_code(dtype),
        int(rows),
        int(cols),
        ptr(out),
        stream_ptr(device),
    )
    return out
