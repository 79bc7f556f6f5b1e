"""Grid-rule restatement of the reference in FP64 numpy (oracle).

Follows torchquad/integration:
  integration_grid.py:64        n = int(N ** (1/dim) + 1e-8)
  integration_grid.py:79-99     per-axis linspace, h = x1 - x0, ij meshgrid points
  simpson.py:54-81, boole.py:56-84   N adjustment
  grid_integrator.py:57-91      reshape [N] -> [n]*dim, then the collapse
  trapezoid.py:28-37, simpson.py:30-45, boole.py:30-48, gaussian.py:132-144
                                per-axis composite collapse of the LAST axis
                                with hs[cur_dim]
  gaussian.py:44-69, :159-162   tensor-product GL weights, resized roots

Pinned by polynomial exactness (trapezoid_test.py:34-48, simpson_test.py
:36-50, boole_test.py:29-33, gauss_test.py:24-45), the tutorial outputs
(tutorial.rst:222, :238, :318, :331, :344), the separable product sums below
(an independent formulation) and the brute-force tensor-product sum of
oracle.c (tests/test_oracle.py).
"""
import functools

import numpy as np


@functools.lru_cache(maxsize=None)
def leggauss(n):
    """numpy's Gauss-Legendre nodes and weights (the reference's _root_fn,
    gaussian.py:25), cached per n; read-only arrays."""
    r, w = np.polynomial.legendre.leggauss(n)
    r.setflags(write=False)
    w.setflags(write=False)
    return r, w


def n_per_dim(N, dim):
    return int(N ** (1.0 / dim) + 1e-8)


def adjust_N(rule, dim, N):
    n = n_per_dim(N, dim)
    if rule == "simpson":
        if n < 3:
            return 3**dim
        if n % 2 != 1:
            return (n - 1) ** dim
    if rule == "boole":
        if n < 5:
            return 5**dim
        if (n - 1) % 4 != 0:
            return (n - ((n - 1) % 4)) ** dim
    return N


def axis_nodes(rule, a, b, n):
    if rule == "gauss_legendre":
        r, _ = leggauss(n)
        return (b - a) / 2.0 * r + (a + b) / 2.0
    return np.linspace(a, b, n)


def grid_points(rule, domain, n):
    domain = np.asarray(domain, dtype=np.float64)
    axes = [axis_nodes(rule, domain[k, 0], domain[k, 1], n) for k in range(domain.shape[0])]
    mesh = np.meshgrid(*axes, indexing="ij")
    return np.stack([m.ravel() for m in mesh], axis=1), axes


def collapse(rule, values_nd, dim, hs, domain):
    """The reference's _apply_composite_rule loops on [*S, n, ..., n]."""
    cur = values_nd
    for cur_dim in range(dim):
        if rule == "trapezoid":
            cur = hs[cur_dim] / 2.0 * (cur[..., 0:-1] + cur[..., 1:])
        elif rule == "simpson":
            cur = hs[cur_dim] / 3.0 * (cur[..., 0:-2][..., ::2] + 4 * cur[..., 1:-1][..., ::2] + cur[..., 2:][..., ::2])
        elif rule == "boole":
            cur = (
                hs[cur_dim]
                / 22.5
                * (
                    7 * cur[..., 0:-4][..., ::4]
                    + 32 * cur[..., 1:-3][..., ::4]
                    + 12 * cur[..., 2:-2][..., ::4]
                    + 32 * cur[..., 3:-1][..., ::4]
                    + 7 * cur[..., 4:][..., ::4]
                )
            )
        elif rule == "gauss_legendre":
            cur = 0.5 * (domain[cur_dim][1] - domain[cur_dim][0]) * cur
        cur = np.sum(cur, axis=-1)
    return cur


def gl_weights(n, dim):
    _, w = leggauss(n)
    mesh = np.meshgrid(*([w] * dim), indexing="ij")
    return np.prod(np.stack(mesh), axis=0).ravel()


def integrate(rule, f, dim, N, domain):
    """Full restatement of Rule().integrate(f, dim, N, domain) in FP64.

    f maps [P, dim] float64 points to [P] or [P, *S] values.
    """
    domain = np.asarray(domain, dtype=np.float64)
    N = adjust_N(rule, dim, N)
    n = n_per_dim(N, dim)
    pts, axes = grid_points(rule, domain, n)
    hs = np.array([ax[1] - ax[0] for ax in axes]) if n > 1 else np.zeros(dim)
    vals = np.asarray(f(pts), dtype=np.float64)
    if rule == "gauss_legendre":
        w = gl_weights(n, dim)
        vals = vals * (w if vals.ndim == 1 else w.reshape((-1,) + (1,) * (vals.ndim - 1)))
    if vals.ndim == 1:
        vals = vals[:, None]
    S = vals.shape[1:]
    vals_nd = np.moveaxis(vals, 0, -1).reshape((*S, *([n] * dim)))
    res = collapse(rule, vals_nd, dim, hs, domain)
    return np.squeeze(res)


def axis_weights(rule, a, b, n):
    """Full per-axis weights (scale included): the 1-D rule."""
    if rule == "gauss_legendre":
        _, w = leggauss(n)
        return w * (b - a) / 2.0
    h = (b - a) / (n - 1)
    w = np.zeros(n)
    if rule == "trapezoid":
        w[:] = 2.0
        w[0] = w[-1] = 1.0
        return w * h / 2.0
    if rule == "simpson":
        w[1::2] = 4.0
        w[2:-1:2] = 2.0
        w[0] = w[-1] = 1.0
        return w * h / 3.0
    if rule == "boole":
        w[1::2] = 32.0
        w[2::4] = 12.0
        w[4:-1:4] = 14.0
        w[0] = w[-1] = 7.0
        return w * h / 22.5
    raise ValueError(rule)


def separable_product(rule, factors, domain, n):
    """prod_k sum_j w_kj g_k(x_kj) for product integrands f = prod_k g_k(x_k)."""
    domain = np.asarray(domain, dtype=np.float64)
    val = 1.0
    for k, g in enumerate(factors):
        a, b = domain[k]
        x = axis_nodes(rule, a, b, n)
        val *= float(np.sum(axis_weights(rule, a, b, n) * g(x)))
    return val


def separable_gaussian_peak(rule, a, u, domain, n):
    return separable_product(rule, [lambda x, ak=ak, uk=uk: np.exp(-((ak * (x - uk)) ** 2)) for ak, uk in zip(a, u)], domain, n)


def separable_oscillatory(rule, c, w, domain, n):
    """Re[e^{i 2 pi w} prod_k sum_j w_kj e^{i c_k x_kj}] (exact identity for cos of a sum)."""
    domain = np.asarray(domain, dtype=np.float64)
    z = np.exp(1j * 2 * np.pi * w)
    for k, ck in enumerate(c):
        a, b = domain[k]
        x = axis_nodes(rule, a, b, n)
        z *= np.sum(axis_weights(rule, a, b, n) * np.exp(1j * ck * x))
    return float(z.real)
