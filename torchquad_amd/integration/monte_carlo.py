"""Monte Carlo integration (reference: torchquad/integration/monte_carlo.py:14-225).

Built-in integrands run as one fused kernel: Philox sampling, the integrand
and a block reduction of (sum f, sum f^2) -- no sample is stored.  The
standard error of the last fused call is kept in ``last_std_error`` (the
reference returns no error estimate).  Any other callable gets the same
sample stream as ``[N, dim]`` points from the HIP point writer, and the mean
is taken by a HIP reduction.
"""
import torch

from .. import _ops
from ..integrands import is_builtin
from ..utils.set_log_level import logger
from .base_integrator import BaseIntegrator
from .rng import RNG
from .utils import _float_domain, _setup_integration_domain, _warn_scalar_deprecation, expand_func_values_and_squeeze_integral


class _McReduceFn(torch.autograd.Function):
    """volume * mean(values) with d/d(values, domain)."""

    @staticmethod
    def forward(ctx, values, domain):
        out = _ops.mc_reduce(values.detach(), domain.detach())
        ctx.save_for_backward(domain, out)
        ctx.meta = (values.shape, values.dtype)
        return out

    @staticmethod
    def backward(ctx, grad_out):
        domain, out = ctx.saved_tensors
        shape, vdtype = ctx.meta
        grad_values = grad_domain = None
        width = (domain[:, 1] - domain[:, 0]).detach()
        vol = torch.prod(width)
        if ctx.needs_input_grad[0]:
            g = (grad_out * (vol / shape[0])).to(vdtype)
            grad_values = g.unsqueeze(0).expand(shape)
        if ctx.needs_input_grad[1]:
            gs = torch.sum(grad_out * out)
            gw = gs / width
            grad_domain = torch.stack([-gw, gw], dim=1).to(domain.dtype)
        return grad_values, grad_domain


class MonteCarlo(BaseIntegrator):
    """Monte Carlo integration"""

    def __init__(self):
        super().__init__()
        self.last_std_error = None
        self._last_stats = None

    def integrate(self, fn, dim, N=1000, integration_domain=None, seed=None, rng=None, backend=None):
        """Integrate fn with plain Monte Carlo (monte_carlo.py:20-58).

        Args:
            fn (func): integrand (built-in family -> fused kernel).
            dim (int): dimensionality of the domain.
            N (int, optional): number of samples. Defaults to 1000.
            integration_domain (list or tensor, optional): defaults to [-1,1]^dim.
            seed (int, optional): Philox key; random when None.
            rng (RNG, optional): an initialised RNG (mutually exclusive with seed).
            backend (string, optional): "torch".

        Returns:
            tensor: integral value
        """
        self._check_inputs(dim=dim, N=N, integration_domain=integration_domain)
        logger.debug("Monte Carlo integrating a %d-dimensional fn with %d points", dim, N)
        integration_domain = _setup_integration_domain(dim, integration_domain, backend)
        domain = _float_domain(integration_domain)
        needs_grad = torch.is_grad_enabled() and domain.requires_grad
        if is_builtin(fn, dim) and not needs_grad:
            if rng is None:
                rng = RNG(backend="torch", seed=seed)
            elif seed is not None:
                raise ValueError("seed and rng cannot both be passed")
            first = rng.reserve(N)
            stats = torch.empty(3, dtype=torch.float64, device=domain.device)
            out = _ops.mc_integrate(fn, dim, N, domain, rng.key, first, stats=stats)
            self._last_stats = stats
            self.last_std_error = None
            self._nr_of_fevals = N
            _warn_scalar_deprecation()
            return out
        sample_points = self.calculate_sample_points(N, integration_domain, seed, rng)
        logger.debug("Evaluating integrand")
        function_values, self._nr_of_fevals = self.evaluate_integrand(fn, sample_points)
        return self.calculate_result(function_values, integration_domain)

    def std_error(self):
        """Standard error of the last fused integral (one device read)."""
        if self._last_stats is None:
            return None
        s1, s2, n = self._last_stats.tolist()
        mean = s1 / n
        var = max(s2 / n - mean * mean, 0.0)
        self.last_std_error = (var / max(n - 1.0, 1.0)) ** 0.5
        return self.last_std_error

    @expand_func_values_and_squeeze_integral
    def calculate_result(self, function_values, integration_domain):
        """volume * mean(values) (monte_carlo.py:60-82), HIP reduction."""
        domain = _float_domain(integration_domain)
        if torch.is_grad_enabled() and (function_values.requires_grad or domain.requires_grad):
            return _McReduceFn.apply(function_values, domain)
        return _ops.mc_reduce(function_values, domain)

    def calculate_sample_points(self, N, integration_domain, seed=None, rng=None):
        """Uniform sample points in the domain (monte_carlo.py:84-106)."""
        if rng is None:
            rng = RNG(backend="torch", seed=seed)
        elif seed is not None:
            raise ValueError("seed and rng cannot both be passed")
        logger.debug("Picking random sampling points")
        domain = _float_domain(integration_domain)
        dim = domain.shape[0]
        if torch.is_grad_enabled() and domain.requires_grad:
            u = rng.uniform(size=[N, dim], dtype=domain.dtype, device=domain.device)
            domain_starts = domain[:, 0]
            domain_sizes = domain[:, 1] - domain_starts
            return u * domain_sizes + domain_starts
        first = rng.reserve(N)
        return _ops.mc_points(dim, N, domain, rng.key, first)

    def get_jit_compiled_integrate(self, dim, N=1000, integration_domain=None, seed=None, backend=None):
        """Integrate function with dim and N fixed (monte_carlo.py:108-225)."""
        integration_domain = _setup_integration_domain(dim, integration_domain, backend)
        self._check_inputs(dim=dim, N=N, integration_domain=integration_domain)
        rng = RNG(backend="torch", seed=seed)

        def compiled_integrate(fn, integration_domain):
            return self.integrate(fn, dim, N, integration_domain, rng=rng, backend=backend)

        return compiled_integrate
