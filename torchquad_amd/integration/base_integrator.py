"""Integrator base class (reference: torchquad/integration/base_integrator.py:9-117)."""
import warnings

import torch

from ..utils.set_log_level import logger
from .utils import _check_integration_domain


class BaseIntegrator:
    """The (abstract) integrator that all other integrators inherit from."""

    _fn = None
    _dim = None
    _integration_domain = None
    _nr_of_fevals = None

    def __init__(self):
        self._nr_of_fevals = 0

    def integrate(self):
        raise (NotImplementedError("This is an abstract base class. Should not be called."))

    def _eval(self, points, weights=None, args=None):
        """evaluate_integrand on self._fn and count evaluations (base_integrator.py:30-40)."""
        result, num_points = self.evaluate_integrand(self._fn, points, weights=weights, args=args)
        self._nr_of_fevals += num_points
        return result

    @staticmethod
    def evaluate_integrand(fn, points, weights=None, args=None):
        """Call ``fn(points, *args)``, check its output and apply weights
        (base_integrator.py:43-91).  Returns (values, number of points)."""
        num_points = points.shape[0]
        if args is None:
            args = ()
        result = fn(points, *args)
        if not isinstance(result, torch.Tensor):
            warnings.warn(
                "The passed function's return value has a different numerical backend than the passed points. "
                "Will try to convert. Note that this may be slow as it results in memory transfers between CPU "
                "and GPU, if torchquad uses the GPU."
            )
            result = torch.as_tensor(result, device=points.device)
        num_results = result.shape[0] if result.dim() > 0 else 1
        if num_results != num_points:
            raise ValueError(
                f"The passed function was given {num_points} points but only returned {num_results} value(s)."
                f"Please ensure that your function is vectorized, i.e. can be called with multiple evaluation "
                f"points at once. It should return a tensor where first dimension matches length of passed "
                f"elements. "
            )
        if weights is not None:
            if len(result.shape) > 1:
                integrand_shape = list(result.shape[1:])
                numel = 1
                for s in integrand_shape:
                    numel *= s
                weights = torch.repeat_interleave(torch.unsqueeze(weights, 1), numel).reshape(
                    (weights.shape[0], *integrand_shape)
                )
            result *= weights
        return result, num_points

    @staticmethod
    def _check_inputs(dim=None, N=None, integration_domain=None):
        """Input validation (base_integrator.py:94-117)."""
        logger.debug("Checking inputs to Integrator.")
        if dim is not None:
            if dim < 1:
                raise ValueError("Dimension needs to be 1 or larger.")
        if N is not None:
            if N < 1 or type(N) is not int:
                raise ValueError("N has to be a positive integer.")
        if integration_domain is not None:
            dim_domain = _check_integration_domain(integration_domain)
            if dim is not None and dim != dim_domain:
                raise ValueError(
                    "The dimension of the integration domain must match the passed function dimensionality dim."
                )
