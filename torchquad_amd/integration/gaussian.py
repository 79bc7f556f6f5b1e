"""Gaussian quadrature (reference: torchquad/integration/gaussian.py:6-162).

``GaussLegendre`` with the default root function runs on the fused HIP
kernels; its node/weight tables live on the device, cached per (n, device).
Subclasses that change ``_root_fn``, ``_resize_roots`` or the composite rule
(e.g. Gauss-Hermite, tutorial.rst:1198-1236) run their own code on the
reference's data layout.

Nodes and weights: ``numpy.polynomial.legendre.leggauss`` (the reference's
``_root_fn``) for n <= ``LEGGAUSS_MAX_N``; above that the package's O(n)
device generator (Newton on asymptotic/recurrence evaluations of P_n), since
leggauss builds a dense n x n companion matrix.
"""
import numpy
import torch

from .. import _ops
from .grid_integrator import GridIntegrator, _composite_by_kernel, _mark_owner

LEGGAUSS_MAX_N = 256


class Gaussian(GridIntegrator):
    """Base class for Gaussian quadrature (Gauss-Legendre on [-1,1] by default).

    The primary hooks are ``_root_fn`` (returns roots and weights like
    ``numpy.polynomial.legendre.leggauss``), ``_resize_roots`` and
    ``_apply_composite_rule``.
    """

    def __init__(self):
        super().__init__()
        self.name = "Gauss-Legendre"
        self._root_fn = numpy.polynomial.legendre.leggauss
        self._root_args = ()
        self._cache = {}
        self._device_cache = {}

    def integrate(self, fn, dim, N=8, integration_domain=None, backend=None):
        """Integrate fn with a Gaussian rule; N is the total number of points (gaussian.py:29-42)."""
        return super().integrate(fn, dim, N, integration_domain, backend)

    def _weights(self, N, dim, backend, requires_grad=False):
        """Tensor-product weights [N^dim] (gaussian.py:44-69)."""
        weights = torch.tensor(self._cached_points_and_weights(N)[1])
        if self._device_cache:
            weights = weights.to(next(iter(self._device_cache.values()))[1].device)
        elif torch.cuda.is_available():
            weights = weights.cuda()
        weights.requires_grad = requires_grad
        mesh = torch.meshgrid(*([weights] * dim), indexing="ij")
        return torch.prod(torch.stack(list(mesh), dim=0), dim=0).ravel()

    def _roots(self, N, backend, requires_grad=False):
        """Roots of the polynomial (gaussian.py:71-84)."""
        roots = torch.tensor(self._cached_points_and_weights(N)[0])
        if torch.cuda.is_available():
            roots = roots.cuda()
        if requires_grad:
            roots.requires_grad = True
        return roots

    @property
    def _grid_func(self):
        """Roots resized to the domain (gaussian.py:86-95)."""

        def f(integration_domain, N, requires_grad, backend=None):
            return self._resize_roots(integration_domain, self._roots(N, backend, requires_grad))

        return f

    def _resize_roots(self, integration_domain, roots):
        """Identity for the base class (gaussian.py:97-107)."""
        return roots

    def _cached_points_and_weights(self, N):
        """Cached roots and weights (gaussian.py:110-129)."""
        _root_args = (N, *self._root_args)
        if not isinstance(N, int):
            if hasattr(N, "item"):
                _root_args = (N.item(), *self._root_args)
            else:
                raise NotImplementedError(f"N {N} is not an int and lacks an `item` method")
        if _root_args in self._cache:
            return self._cache[_root_args]
        if self._root_fn is numpy.polynomial.legendre.leggauss and _root_args[0] > LEGGAUSS_MAX_N:
            from ..gl_nodes import gauss_legendre

            r, w = gauss_legendre(_root_args[0])
            self._cache[_root_args] = (r.cpu().numpy(), w.cpu().numpy())
        else:
            self._cache[_root_args] = self._root_fn(*_root_args)
        return self._cache[_root_args]

    @staticmethod
    def _apply_composite_rule(cur_dim_areas, dim, hs, domain):
        """Sum with (b-a)/2 per axis (gaussian.py:132-144)."""
        for cur_dim in range(dim):
            cur_dim_areas = (
                0.5 * (domain[cur_dim][1] - domain[cur_dim][0]) * torch.sum(cur_dim_areas, dim=len(cur_dim_areas.shape) - 1)
            )
        return cur_dim_areas


@_mark_owner
class GaussLegendre(Gaussian):
    """Gauss-Legendre quadrature on any box: x = (b-a)/2 r + (a+b)/2 (gaussian.py:147-162)."""

    _tq_rule = _ops.RULES["gauss_legendre"]

    def __init__(self):
        super().__init__()

    def _tq_standard(self):
        return (
            super()._tq_standard()
            and type(self)._resize_roots is GaussLegendre._resize_roots
            and type(self)._roots is Gaussian._roots
            and type(self)._cached_points_and_weights is Gaussian._cached_points_and_weights
            and self._root_fn is numpy.polynomial.legendre.leggauss
            and self._root_args == ()
        )

    def _gl_tables(self, n, device):
        """Device float64 (nodes, weights) on [-1, 1], cached per (n, device)."""
        key = (int(n), str(device))
        t = self._device_cache.get(key)
        if t is None:
            if self._root_fn is numpy.polynomial.legendre.leggauss and n > LEGGAUSS_MAX_N:
                from ..gl_nodes import gauss_legendre

                t = gauss_legendre(int(n), device=device)
            else:
                r, w = self._cached_points_and_weights(int(n))
                t = (
                    torch.tensor(numpy.asarray(r, dtype=numpy.float64), device=device),
                    torch.tensor(numpy.asarray(w, dtype=numpy.float64), device=device),
                )
            self._device_cache[key] = t
        return t

    @property
    def _grid_func(self):
        f = Gaussian._grid_func.fget(self)
        if self._tq_standard():
            f._tq_spec = ("gauss_legendre", self._gl_tables)
        return f

    def _resize_roots(self, integration_domain, roots):
        a = integration_domain[0]
        b = integration_domain[1]
        return ((b - a) / 2) * roots + ((a + b) / 2)

    @staticmethod
    def _apply_composite_rule(cur_dim_areas, dim, hs, domain):
        """Sum with (b-a)/2 per axis via the HIP reduction (gaussian.py:132-144)."""
        return _composite_by_kernel(_ops.RULES["gauss_legendre"], cur_dim_areas, dim, hs, domain)


@_mark_owner
class GaussHermite(Gaussian):
    """Gauss-Hermite quadrature: integrals of exp(-|x|^2) f(x) over R^dim.

    The reference shows this rule as a user subclass of Gaussian
    (tutorial.rst:1198-1236): roots and weights from
    ``scipy.special.roots_hermite``, identity ``_resize_roots`` and a plain
    sum as composite rule, so the integration domain only sets dtype and
    device.  Here it runs on the fused and point/reduce kernels of the
    Gauss-Legendre path with the Hermite node/weight tables on the unit box
    [-1, 1]^dim, where the affine map is the identity and the per-axis scale
    (b - a)/2 is 1.  A subclass that overrides a hook runs its own code on
    the reference's layout.
    """

    _tq_rule = _ops.RULES["gauss_legendre"]

    def __init__(self):
        super().__init__()
        import scipy.special

        self.name = "Gauss-Hermite"
        self._root_fn = scipy.special.roots_hermite

    def integrate(self, fn, dim, N=8, integration_domain=None, backend=None):
        """Integrate exp(-|x|^2) fn(x) over R^dim with N total points."""
        from .utils import _float_domain, _setup_integration_domain

        dom = _float_domain(_setup_integration_domain(dim, integration_domain, backend))
        unit = torch.tensor([[-1.0, 1.0]] * dim, dtype=dom.dtype, device=dom.device)
        return super().integrate(fn, dim, N, unit, backend)

    def _tq_standard(self):
        import scipy.special

        return (
            super()._tq_standard()
            and type(self)._resize_roots is Gaussian._resize_roots
            and type(self)._roots is Gaussian._roots
            and type(self)._cached_points_and_weights is Gaussian._cached_points_and_weights
            and self._root_fn is scipy.special.roots_hermite
            and self._root_args == ()
        )

    def _gl_tables(self, n, device):
        """Device float64 (nodes, weights) of the n-point Hermite rule, cached per (n, device)."""
        key = (int(n), str(device))
        t = self._device_cache.get(key)
        if t is None:
            r, w = self._cached_points_and_weights(int(n))
            t = (
                torch.tensor(numpy.asarray(r, dtype=numpy.float64), device=device),
                torch.tensor(numpy.asarray(w, dtype=numpy.float64), device=device),
            )
            self._device_cache[key] = t
        return t

    @property
    def _grid_func(self):
        f = Gaussian._grid_func.fget(self)
        if self._tq_standard():
            f._tq_spec = ("gauss_legendre", self._gl_tables)
        return f

    @staticmethod
    def _apply_composite_rule(cur_dim_areas, dim, hs, domain):
        """Plain sum over each axis (the tutorial's Gauss-Hermite composite rule)."""
        for _ in range(dim):
            cur_dim_areas = torch.sum(cur_dim_areas, dim=len(cur_dim_areas.shape) - 1)
        return cur_dim_areas
