"""Trapezoidal rule (reference: torchquad/integration/trapezoid.py:6-37)."""
from .. import _ops
from .grid_integrator import _composite_by_kernel, _mark_owner
from .newton_cotes import NewtonCotes


@_mark_owner
class Trapezoid(NewtonCotes):
    """Trapezoidal rule; per-axis weights h/2*[1, 2, ..., 2, 1]."""

    _tq_rule = _ops.RULES["trapezoid"]

    def __init__(self):
        super().__init__()

    def integrate(self, fn, dim, N=1000, integration_domain=None, backend=None):
        """Integrate fn over the domain with the trapezoid rule (trapezoid.py:12-26).

        Args:
            fn (func): integrand; a built-in family from torchquad_amd.integrands
                runs fused on the GPU, any other callable gets [N, dim] points.
            dim (int): dimensionality of the integration domain.
            N (int, optional): total number of points. Defaults to 1000.
            integration_domain (list or tensor, optional): defaults to [-1,1]^dim.
            backend (string, optional): "torch".

        Returns:
            tensor: integral value
        """
        return super().integrate(fn, dim, N, integration_domain, backend)

    @staticmethod
    def _apply_composite_rule(cur_dim_areas, dim, hs, domain):
        """Composite trapezoid collapse of [*S, n, ..., n] (trapezoid.py:28-37)."""
        return _composite_by_kernel(_ops.RULES["trapezoid"], cur_dim_areas, dim, hs, domain)
