"""Simpson's rule (reference: torchquad/integration/simpson.py:8-81)."""
import warnings

from .. import _ops
from ..utils.set_log_level import logger
from .grid_integrator import _composite_by_kernel, _mark_owner
from .newton_cotes import NewtonCotes


@_mark_owner
class Simpson(NewtonCotes):
    """Simpson's rule; per-axis weights h/3*[1, 4, 2, 4, ..., 2, 4, 1], odd n >= 3."""

    _tq_rule = _ops.RULES["simpson"]

    def __init__(self):
        super().__init__()

    def integrate(self, fn, dim, N=None, integration_domain=None, backend=None):
        """Integrate fn with Simpson's rule; N defaults to 3^dim (simpson.py:14-28)."""
        return super().integrate(fn, dim, N, integration_domain, backend)

    @staticmethod
    def _apply_composite_rule(cur_dim_areas, dim, hs, domain):
        """Composite Simpson collapse of [*S, n, ..., n] (simpson.py:30-45)."""
        return _composite_by_kernel(_ops.RULES["simpson"], cur_dim_areas, dim, hs, domain)

    @staticmethod
    def _get_minimal_N(dim):
        """simpson.py:49-51"""
        return 3**dim

    @staticmethod
    def _adjust_N(dim, N):
        """Lower N to an odd number of points per axis >= 3 (simpson.py:54-81)."""
        n_per_dim = int(N ** (1.0 / dim) + 1e-8)
        logger.debug("Checking if N per dim is >=3 and odd.")
        if n_per_dim < 3:
            warnings.warn("N per dimension cannot be lower than 3. " "N per dim will now be changed to 3.")
            N = 3**dim
        elif n_per_dim % 2 != 1:
            warnings.warn(
                "N per dimension cannot be even due to necessary subdivisions. "
                "N per dim will now be changed to the next lower integer, i.e. "
                f"{n_per_dim} -> {n_per_dim - 1}."
            )
            N = (n_per_dim - 1) ** (dim)
        return N
