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
