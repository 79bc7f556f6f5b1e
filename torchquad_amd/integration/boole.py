"""Boole's rule (reference: torchquad/integration/boole.py:8-84)."""
import warnings

from .. import _ops
from ..utils.set_log_level import logger
from .grid_integrator import _composite_by_kernel, _mark_owner
from .newton_cotes import NewtonCotes


@_mark_owner
class Boole(NewtonCotes):
    """Boole's rule; per-axis weights h/22.5*[7, 32, 12, 32, 14, ..., 32, 7], n = 1 + 4k >= 5."""

    _tq_rule = _ops.RULES["boole"]

    def __init__(self):
        super().__init__()

    def integrate(self, fn, dim, N=None, integration_domain=None, backend=None):
        """Integrate fn with Boole's rule; N defaults to 5^dim (boole.py:14-28)."""
        return super().integrate(fn, dim, N, integration_domain, backend)

    @staticmethod
    def _apply_composite_rule(cur_dim_areas, dim, hs, domain):
        """Composite Boole collapse of [*S, n, ..., n] (boole.py:30-48)."""
        return _composite_by_kernel(_ops.RULES["boole"], cur_dim_areas, dim, hs, domain)

    @staticmethod
    def _get_minimal_N(dim):
        """boole.py:51-53"""
        return 5**dim

    @staticmethod
    def _adjust_N(dim, N):
        """Lower N so that n per axis = 1 + 4k >= 5 (boole.py:56-84)."""
        n_per_dim = int(N ** (1.0 / dim) + 1e-8)
        logger.debug("Checking if N per dim is >=5 and N = 1 + 4n, where n is a positive integer.")
        if n_per_dim < 5:
            warnings.warn("N per dimension cannot be lower than 5. " "N per dim will now be changed to 5.")
            N = 5**dim
        elif (n_per_dim - 1) % 4 != 0:
            new_n_per_dim = n_per_dim - ((n_per_dim - 1) % 4)
            warnings.warn(
                "N per dimension must be N = 1 + 4n with n a positive integer due to necessary subdivisions. "
                "N per dim will now be changed to the next lower N satisfying this, i.e. "
                f"{n_per_dim} -> {new_n_per_dim}."
            )
            N = (new_n_per_dim) ** (dim)
        return N
