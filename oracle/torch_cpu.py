"""Torch-CPU restatement of the reference's integration path (oracle; used
only as the timed CPU baseline of bench.py).

It runs the reference's algorithm with torch on the host's cores:

* grid rules (integration_grid.py:28-103, grid_integrator.py:57-91,
  boole.py:30-48, trapezoid.py:28-37, simpson.py:30-45): per-axis
  ``torch.linspace``, the ``meshgrid(indexing="ij")`` points, the integrand
  on the ``[N, dim]`` points, then the composite rule collapsing the last
  axis ``dim`` times.  The outermost axis is processed one node at a time
  (each slice is the reference's computation on an ``[n^(dim-1), dim]``
  point set) so that memory stays bounded; the slice values are collapsed
  over that axis by the same rule at the end.  Axis k is scaled by its own
  h_k (the reference indexes hs by collapse step, which is the same for the
  equal widths of every benchmarked domain);
* Monte Carlo (monte_carlo.py:60-106): ``torch.rand`` points, the integrand
  and the volume-scaled mean, in chunks.
"""
import math
import time

import torch


def _collapse(rule, values, h):
    """One application of the reference's composite rule to the last axis."""
    if rule == "trapezoid":
        return h / 2.0 * torch.sum(values[..., 0:-1] + values[..., 1:], dim=-1)
    if rule == "simpson":
        return h / 3.0 * torch.sum(values[..., 0:-2][..., ::2] + 4 * values[..., 1:-1][..., ::2]
                                   + values[..., 2:][..., ::2], dim=-1)
    if rule == "boole":
        return h / 22.5 * torch.sum(7 * values[..., 0:-4][..., ::4] + 32 * values[..., 1:-3][..., ::4]
                                    + 12 * values[..., 2:-2][..., ::4] + 32 * values[..., 3:-1][..., ::4]
                                    + 7 * values[..., 4:][..., ::4], dim=-1)
    raise ValueError(rule)


def grid_rule_slices(rule, fn, domain, n, dtype=torch.float32, seconds=None):
    """The reference's grid rule on torch CPU, one outer-axis slice at a time.

    fn: torch callable on [M, dim] points.  Returns (value or None, slices
    done, points evaluated, seconds): the value only when every slice ran
    (``seconds`` bounds the time spent; at least one slice runs).
    """
    dim = len(domain)
    axes = [torch.linspace(float(lo), float(hi), n, dtype=dtype) for lo, hi in domain]
    hs = [ax[1] - ax[0] for ax in axes]
    inner = torch.meshgrid(*axes[1:], indexing="ij") if dim > 1 else ()
    inner = [m.reshape(-1) for m in inner]
    slice_vals = []
    t0 = time.perf_counter()
    for i in range(n):
        x0 = axes[0][i].expand(inner[0].shape[0] if inner else 1)
        pts = torch.stack([x0, *inner], dim=1)
        vals = fn(pts)
        vals = vals.reshape([n] * (dim - 1)) if dim > 1 else vals.reshape(())
        for k in range(dim - 1, 0, -1):
            vals = _collapse(rule, vals, hs[k])
        slice_vals.append(vals)
        if seconds is not None and time.perf_counter() - t0 >= seconds:
            break
    dt = time.perf_counter() - t0
    done = len(slice_vals)
    value = None
    if done == n:
        value = float(_collapse(rule, torch.stack(slice_vals), hs[0]))
    return value, done, done * n ** (dim - 1), dt


def gaussian_peak_fn(a, u, dtype=torch.float32):
    at = torch.as_tensor(a, dtype=dtype)
    ut = torch.as_tensor(u, dtype=dtype)
    return lambda x: torch.exp(-torch.sum((at * (x - ut)) ** 2, dim=1))


def exp_sum_fn():
    return lambda x: torch.exp(torch.sum(x, dim=1))


def monte_carlo_oscillatory(c, w, domain, seconds, chunk=10_000_000, dtype=torch.float32, seed=0):
    """Plain MC of cos(2 pi w + sum c_k x_k) with torch.rand points, chunked;
    returns (samples, seconds, estimate)."""
    g = torch.Generator().manual_seed(seed)
    dom = torch.tensor(domain, dtype=dtype)
    lo, size = dom[:, 0], dom[:, 1] - dom[:, 0]
    ct = torch.as_tensor(c, dtype=dtype)
    vol = float(torch.prod(size))
    total, acc = 0, 0.0
    t0 = time.perf_counter()
    while True:
        x = torch.rand((chunk, dom.shape[0]), generator=g, dtype=dtype) * size + lo
        f = torch.cos(2.0 * math.pi * w + x @ ct)
        acc += float(torch.sum(f, dtype=torch.float64))
        total += chunk
        if time.perf_counter() - t0 >= seconds:
            break
    return total, time.perf_counter() - t0, vol * acc / total
