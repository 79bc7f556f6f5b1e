"""Gaussian quadrature (reference: torchquad/integration/gaussian.py:6-162).

Any Gaussian rule -- the built-in ``GaussLegendre``, ``GaussHermite`` or a
user subclass that sets ``_root_fn`` (and possibly overrides
``_resize_roots`` / ``_apply_composite_rule``, as the tutorial's
Gauss-Hermite does, tutorial.rst:1198-1236) -- runs on the fused and
point/reduce HIP kernels with its node/weight tables uploaded once per
(n, device, root function).  The kernels implement

    x_k = (b_k - a_k)/2 r + (a_k + b_k)/2,   I = prod_k (b_k - a_k)/2 sum W f,

and the two hooks that shape a rule are classified once per class by
evaluating them on small probe tensors:

* ``_resize_roots``: affine (Gauss-Legendre's map to [a, b]) or identity;
* ``_apply_composite_rule``: the sum scaled by (b - a)/2 per axis (the base
  class's rule) or the plain sum.

(affine, scaled) is the kernels' rule on the given domain; (identity, plain
sum) is the kernels' rule on the unit box [-1, 1]^dim, where both the map
and the scale are the identity.  Other combinations, and classes that
override ``_roots``, ``_weights``, ``_grid_func`` or ``calculate_result``,
run their own code on the reference's data layout.

Nodes and weights come from ``_cached_points_and_weights`` (the reference's
cache around ``_root_fn``), except Gauss-Legendre above ``LEGGAUSS_MAX_N``
nodes, where the package's O(n) device generator replaces
``numpy.polynomial.legendre.leggauss`` (a dense n x n eigenproblem).
"""
import numpy
import torch

from .. import _ops
from .grid_integrator import GridIntegrator

LEGGAUSS_MAX_N = 256

AFFINE = "affine"  # kernels' map and (b-a)/2 scale on the given domain
UNIT = "unit"      # identity map and plain sum: the kernels on [-1, 1]^dim

_mode_cache = {}


def _classify_resize(resize):
    """'identity', 'affine' or None for a _resize_roots bound method."""
    r = torch.tensor([-0.7, 0.1, 0.55], dtype=torch.float64)
    seen = set()
    for a, b in ((0.3, 1.7), (-2.0, -0.5)):
        dom = torch.tensor([a, b], dtype=torch.float64)
        out = torch.as_tensor(resize(dom, r.clone()), dtype=torch.float64)
        if tuple(out.shape) != tuple(r.shape):
            return None
        if torch.allclose(out, r, rtol=1e-14, atol=0.0):
            seen.add("identity")
        elif torch.allclose(out, 0.5 * (b - a) * r + 0.5 * (a + b), rtol=1e-14, atol=1e-15):
            seen.add("affine")
        else:
            return None
    return seen.pop() if len(seen) == 1 else None


def _classify_composite(rule):
    """'sum', 'scaled' or None for an _apply_composite_rule callable."""
    g = torch.Generator().manual_seed(7)
    vals = torch.rand((2, 3, 3), generator=g, dtype=torch.float64) + 0.5
    dom = torch.tensor([[0.3, 1.7], [-1.0, 2.0]], dtype=torch.float64)
    hs = torch.tensor([0.1, 0.2], dtype=torch.float64)
    out = torch.as_tensor(rule(vals.clone(), 2, hs, dom), dtype=torch.float64)
    plain = vals.sum(dim=(-1, -2))
    if tuple(out.shape) != tuple(plain.shape):
        return None
    if torch.allclose(out, plain, rtol=1e-13, atol=0.0):
        return "sum"
    if torch.allclose(out, 0.7 * 1.5 * plain, rtol=1e-13, atol=0.0):
        return "scaled"
    return None


class Gaussian(GridIntegrator):
    """Base class for Gaussian quadrature (Gauss-Legendre roots on [-1,1] by default).

    The primary hooks are ``_root_fn`` (returns roots and weights like
    ``numpy.polynomial.legendre.leggauss``), ``_resize_roots`` and
    ``_apply_composite_rule``.
    """

    _tq_rule = _ops.RULES["gauss_legendre"]

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

    # -- which kernels can run this rule --------------------------------------
    def _tq_mode(self):
        """AFFINE, UNIT or None (see the module docstring); cached per class."""
        cls = type(self)
        if (
            cls._roots is not Gaussian._roots
            or cls._weights is not Gaussian._weights
            or cls._grid_func is not Gaussian._grid_func
            or cls.calculate_result is not GridIntegrator.calculate_result
            or cls._cached_points_and_weights is not Gaussian._cached_points_and_weights
        ):
            return None
        key = (cls, cls.__dict__.get("_resize_roots"), cls.__dict__.get("_apply_composite_rule"))
        if key in _mode_cache:
            return _mode_cache[key]
        try:
            resize = _classify_resize(self._resize_roots)
            comp = _classify_composite(self._apply_composite_rule)
        except Exception:  # a user hook that does not accept the probes: its own code runs
            resize = comp = None
        mode = {("affine", "scaled"): AFFINE, ("identity", "sum"): UNIT}.get((resize, comp))
        _mode_cache[key] = mode
        return mode

    def _tq_standard(self):
        return self._tq_mode() is not None

    def _tq_domain(self, domain):
        """The domain the kernels see: the unit box for identity-map, plain-sum rules."""
        if self._tq_mode() != UNIT:
            return domain
        if isinstance(domain, torch.Tensor):
            unit = torch.empty_like(domain)
            unit[:, 0] = -1.0
            unit[:, 1] = 1.0
            return unit
        return [[-1.0, 1.0] for _ in domain]

    def _gl_tables(self, n, device):
        """Device float64 (nodes, weights) of the n-point rule, cached per
        (n, device, root function)."""
        key = (int(n), str(device), id(self._root_fn), tuple(self._root_args))
        t = self._device_cache.get(key)
        if t is None:
            if self._root_fn is numpy.polynomial.legendre.leggauss and not self._root_args and n > LEGGAUSS_MAX_N:
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

    # -- the reference's hooks -----------------------------------------------
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

        if self._tq_standard():
            f._tq_spec = ("gauss_legendre", self._gl_tables)
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


class GaussLegendre(Gaussian):
    """Gauss-Legendre quadrature on any box: x = (b-a)/2 r + (a+b)/2 (gaussian.py:147-162)."""

    def __init__(self):
        super().__init__()

    def _resize_roots(self, integration_domain, roots):
        a = integration_domain[0]
        b = integration_domain[1]
        return ((b - a) / 2) * roots + ((a + b) / 2)


class GaussHermite(Gaussian):
    """Gauss-Hermite quadrature: integrals of exp(-|x|^2) f(x) over R^dim.

    The tutorial's custom rule (tutorial.rst:1198-1236): roots and weights
    from ``scipy.special.roots_hermite``, identity ``_resize_roots`` and a
    plain sum as composite rule, so the integration domain only sets dtype
    and device.  It runs on the kernels' unit box [-1, 1]^dim like any user
    subclass written the same way.
    """

    def __init__(self):
        super().__init__()
        import scipy.special

        self.name = "Gauss-Hermite"
        self._root_fn = scipy.special.roots_hermite

    @staticmethod
    def _apply_composite_rule(cur_dim_areas, dim, hs, domain):
        """Plain sum over each axis (the tutorial's Gauss-Hermite composite rule)."""
        for _ in range(dim):
            cur_dim_areas = torch.sum(cur_dim_areas, dim=len(cur_dim_areas.shape) - 1)
        return cur_dim_areas
