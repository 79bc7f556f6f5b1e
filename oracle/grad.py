"""Gradient oracles for the Gaussian peak (test infrastructure only).

closed_form(a, u, lo, hi): the integral over the box and its exact
derivatives w.r.t. a, u, lo, hi from the erf closed form
    G_k = sqrt(pi)/(2 a_k) [erf(a_k (hi_k - u_k)) - erf(a_k (lo_k - u_k))],
    I = prod_k G_k.
reference_autograd(...): torch (CPU, float64) restatement of the reference's
GaussLegendre path -- nodes from numpy leggauss, _resize_roots
(gaussian.py:159-162), tensor-product weights (gaussian.py:44-69), the
(b-a)/2 collapse (gaussian.py:132-144) -- differentiated by torch autograd,
which is how the reference obtains gradients (tests/gradient_test.py).
"""
import math

import numpy as np
import torch
from scipy.special import erf


def closed_form(a, u, lo, hi):
    a, u, lo, hi = map(lambda v: np.asarray(v, dtype=np.float64), (a, u, lo, hi))
    A, Bv = hi - u, lo - u
    G = math.sqrt(math.pi) / (2 * a) * (erf(a * A) - erf(a * Bv))
    eA, eB = np.exp(-(a * A) ** 2), np.exp(-(a * Bv) ** 2)
    I = float(np.prod(G))
    dG_dhi = eA
    dG_dlo = -eB
    dG_du = -(eA - eB)
    dG_da = -G / a + (A * eA - Bv * eB) / a
    f = I / G
    return I, f * dG_da, f * dG_du, f * dG_dlo, f * dG_dhi


def reference_autograd(a, u, lo, hi, n):
    a = torch.tensor(np.asarray(a, dtype=np.float64), requires_grad=True)
    u = torch.tensor(np.asarray(u, dtype=np.float64), requires_grad=True)
    dom = torch.tensor(np.stack([lo, hi], axis=1).astype(np.float64), requires_grad=True)
    dim = a.shape[0]
    r, w = np.polynomial.legendre.leggauss(n)
    roots = torch.tensor(r)
    wt = torch.tensor(w)
    axes = [((dom[k, 1] - dom[k, 0]) / 2) * roots + ((dom[k, 0] + dom[k, 1]) / 2) for k in range(dim)]
    mesh = torch.meshgrid(*axes, indexing="ij")
    pts = torch.stack([m.ravel() for m in mesh], dim=1)
    vals = torch.exp(-torch.sum((a * (pts - u)) ** 2, dim=1))
    W = torch.prod(torch.stack(list(torch.meshgrid(*([wt] * dim), indexing="ij"))), dim=0).ravel()
    vals = vals * W
    cur = vals.reshape([n] * dim)
    for k in range(dim):
        cur = 0.5 * (dom[k][1] - dom[k][0]) * torch.sum(cur, dim=-1)
    cur.backward()
    return float(cur), a.grad.numpy(), u.grad.numpy(), dom.grad[:, 0].numpy(), dom.grad[:, 1].numpy()


def _torch_family(family, pts, p1, p2):
    """The family's integrand on [M, dim] points (torch, differentiable)."""
    if family == "gaussian_peak":
        return torch.exp(-torch.sum((p1 * (pts - p2)) ** 2, dim=1))
    if family == "oscillatory":
        return torch.cos(2 * math.pi * p2 + pts @ p1)
    if family == "product_peak":
        return torch.prod(1.0 / (p1 ** -2 + (pts - p2) ** 2), dim=1)
    raise ValueError(family)


def reference_autograd_family(family, p1, p2, lo, hi, n, rule="gauss_legendre"):
    """The reference's grid path for one integrand, in torch FP64 on the CPU,
    differentiated by torch autograd (tests/gradient_test.py:162-259).

    Gauss-Legendre: roots resized per axis (gaussian.py:159-162) with the
    tensor-product weights (gaussian.py:44-69) and the (b-a)/2 collapse
    (gaussian.py:132-144).  Newton-Cotes: linspace nodes whose positions
    depend on the bounds (integration_grid.py:28-103, utils.py:24-58) and
    the composite collapse with h = (hi - lo)/(n - 1) (trapezoid.py:28-37,
    simpson.py:30-45, boole.py:30-48).  Parameters: Gaussian and product peak
    a = p1 [dim], u = p2 [dim]; oscillatory c = p1 [dim], w = p2 (scalar).
    Returns (I, grad p1, grad p2, grad lo, grad hi)."""
    t1 = torch.tensor(np.asarray(p1, dtype=np.float64), requires_grad=True)
    t2 = torch.tensor(np.asarray(p2, dtype=np.float64), requires_grad=True)
    dom = torch.tensor(np.stack([lo, hi], axis=1).astype(np.float64), requires_grad=True)
    dim = t1.shape[0]
    if rule == "gauss_legendre":
        r, w = np.polynomial.legendre.leggauss(n)
        roots, wt = torch.tensor(r), torch.tensor(w)
        axes = [((dom[k, 1] - dom[k, 0]) / 2) * roots + ((dom[k, 0] + dom[k, 1]) / 2) for k in range(dim)]
    else:
        frac = torch.linspace(0.0, 1.0, n, dtype=torch.float64)
        axes = [dom[k, 0] + (dom[k, 1] - dom[k, 0]) * frac for k in range(dim)]
    mesh = torch.meshgrid(*axes, indexing="ij")
    pts = torch.stack([m.ravel() for m in mesh], dim=1)
    vals = _torch_family(family, pts, t1, t2)
    cur = vals.reshape([n] * dim)
    for k in range(dim):
        if rule == "gauss_legendre":
            cur = 0.5 * (dom[k][1] - dom[k][0]) * torch.sum(cur * wt, dim=-1)
        else:
            h = (dom[k][1] - dom[k][0]) / (n - 1)
            wk = torch.tensor(_nc_unit_weights(rule, n))
            cur = h * torch.sum(cur * wk, dim=-1)
    cur.backward()
    return (float(cur.detach()), t1.grad.numpy(), np.asarray(t2.grad.numpy()), dom.grad[:, 0].numpy(),
            dom.grad[:, 1].numpy())


def _nc_unit_weights(rule, n):
    """Composite Newton-Cotes weights for unit spacing (h factored out)."""
    w = np.zeros(n)
    if rule == "trapezoid":
        w[:] = 1.0
        w[0] = w[-1] = 0.5
    elif rule == "simpson":
        w[1::2] = 4.0
        w[2:-1:2] = 2.0
        w[0] = w[-1] = 1.0
        w /= 3.0
    elif rule == "boole":
        w[1::2] = 32.0
        w[2::4] = 12.0
        w[4:-1:4] = 14.0
        w[0] = w[-1] = 7.0
        w /= 22.5
    else:
        raise ValueError(rule)
    return w
