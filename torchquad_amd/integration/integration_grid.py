"""Integration grid (reference: torchquad/integration/integration_grid.py:19-124).

The standard grids (``linspace`` per axis for Newton-Cotes rules, resized
Gauss-Legendre roots) are written by the HIP point writer; gradients with
respect to the domain flow through a hand-written backward kernel.  Custom
``grid_func`` callables (e.g. subclasses overriding ``_grid_func``) are
called per axis and combined with ``meshgrid`` exactly as in the reference.
"""
from time import perf_counter

import torch

from .. import _ops
from ..utils.set_log_level import logger
from .utils import _check_integration_domain, _linspace_with_grads, _setup_integration_domain


def grid_func(integration_domain, N, requires_grad=False, backend=None):
    a = integration_domain[0]
    b = integration_domain[1]
    return _linspace_with_grads(a, b, N, requires_grad=requires_grad)


# Marks the standard grid functions: (rule code, gl table provider or None).
grid_func._tq_spec = ("linspace", None)


class _GridPointsFn(torch.autograd.Function):
    """points = grid(domain) with d points / d domain (utils.py:24-58, gaussian.py:159-162)."""

    @staticmethod
    def forward(ctx, domain, rule, dim, n, gl):
        ctx.meta = (rule, dim, n, gl)
        ctx.save_for_backward(domain)
        return _ops.grid_points(rule, dim, n, domain.detach(), gl)

    @staticmethod
    def backward(ctx, grad_points):
        (domain,) = ctx.saved_tensors
        rule, dim, n, gl = ctx.meta
        gd = _ops.grid_points_backward(rule, dim, n, domain.detach(), grad_points, gl)
        return gd, None, None, None, None


def standard_points(rule, dim, n, domain, gl=None):
    if domain.requires_grad and torch.is_grad_enabled():
        return _GridPointsFn.apply(domain, rule, dim, n, gl)
    return _ops.grid_points(rule, dim, n, domain, gl)


def standard_h(rule, dim, n, domain, gl=None):
    """h_d = x_1 - x_0 (integration_grid.py:90-93)."""
    if domain.requires_grad and torch.is_grad_enabled():
        width = domain[:, 1] - domain[:, 0]
        if rule == _ops.RULES["gauss_legendre"]:
            r = gl[0]
            return width * (0.5 * (r[1] - r[0]).to(domain.dtype)) if n > 1 else width * 0.0
        return width / (n - 1)
    return _ops.grid_h(rule, dim, n, domain, gl)


class IntegrationGrid:
    """Stores the integration grid for grid rules (integration_grid.py:19-124)."""

    points = None
    h = None
    _N = None
    _dim = None
    _runtime = None

    def __init__(self, N, integration_domain, grid_func=grid_func, disable_integration_domain_check=False):
        start = perf_counter()
        self._check_inputs(N, integration_domain, disable_integration_domain_check)
        if not isinstance(integration_domain, torch.Tensor):
            integration_domain = _setup_integration_domain(len(integration_domain), integration_domain, backend="torch")
        elif not integration_domain.dtype.is_floating_point:
            integration_domain = integration_domain.to(torch.float64)

        self._dim = integration_domain.shape[0]
        # int() floors, so add a small term: (4**3)**(1/3) = 3.99999...
        self._N = int(N ** (1.0 / self._dim) + 1e-8)

        requires_grad = bool(getattr(integration_domain, "requires_grad", False))
        spec = getattr(grid_func, "_tq_spec", None)
        standard = (
            spec is not None
            and integration_domain.dim() == 2
            and integration_domain.shape[1] == 2
            and integration_domain.device.type == "cuda"
        )
        if standard:
            kind, gl_provider = spec
            dom = integration_domain if requires_grad else integration_domain.detach()
            dom = dom.contiguous()
            if kind == "linspace":
                rule = _ops.RULES["trapezoid"]
                gl = None
            else:
                rule = _ops.RULES["gauss_legendre"]
                gl = gl_provider(self._N, dom.device)
            self.h = standard_h(rule, self._dim, self._N, dom, gl)
            self.points = standard_points(rule, self._dim, self._N, dom, gl)
        else:
            grid_1d = [
                grid_func(integration_domain[d], self._N, requires_grad=requires_grad, backend="torch")
                for d in range(self._dim)
            ]
            self.h = torch.stack([grid_1d[d][1] - grid_1d[d][0] for d in range(self._dim)])
            mesh = torch.meshgrid(*grid_1d, indexing="ij")
            self.points = torch.stack([mg.ravel() for mg in mesh], dim=1)
        logger.info("Integration grid created.")
        self._runtime = perf_counter() - start

    def _check_inputs(self, N, integration_domain, disable_integration_domain_check):
        """integration_grid.py:105-124"""
        logger.debug("Checking inputs to IntegrationGrid.")
        if disable_integration_domain_check:
            dim = len(integration_domain)
        else:
            dim = _check_integration_domain(integration_domain)
        if N < 2:
            raise ValueError("N has to be > 1.")
        if N ** (1.0 / dim) < 2:
            raise ValueError(
                "Cannot create a ",
                dim,
                "-dimensional grid with ",
                N,
                " points. Too few points per dimension.",
            )
