"""Monte Carlo integration (reference: torchquad/integration/monte_carlo.py:14-225).

Built-in integrands run as one fused kernel: Philox sampling, the integrand
and a block reduction of (sum f, sum f^2) -- no sample is stored.  The
standard error of the last fused call is kept in ``last_std_error`` (the
reference returns no error estimate).  Any other callable gets the same
sample stream as ``[N, dim]`` points from the HIP point writer, and the mean
is taken by a HIP reduction.
"""
import warnings

import numpy as np
import torch

from .. import _ops
from ..integrands import is_batched, is_builtin
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
            self.last_std_error = None
            self._nr_of_fevals = N
            if is_batched(fn):
                return self._param_batch_integrate(fn, dim, N, domain, rng.key, first)
            stats = torch.empty(3, dtype=torch.float64, device=domain.device)
            out = _ops.mc_integrate(fn, dim, N, domain, rng.key, first, stats=stats)
            self._last_stats = stats
            _warn_scalar_deprecation()
            return out
        sample_points = self.calculate_sample_points(N, integration_domain, seed, rng)
        logger.debug("Evaluating integrand")
        function_values, self._nr_of_fevals = self.evaluate_integrand(fn, sample_points)
        return self.calculate_result(function_values, integration_domain)

    def _param_batch_integrate(self, fn, dim, N, domain, key, first):
        """[*S] integrals of a built-in family's parameter batch, every set on
        the same samples (as an [N, *S] integrand sees one set of points,
        monte_carlo.py:60-106): one fused launch per set."""
        descs = fn.descriptors(dim)
        out = torch.empty(len(descs), dtype=domain.dtype, device=domain.device)
        stats = torch.empty((len(descs), 3), dtype=torch.float64, device=domain.device)
        for i, d in enumerate(descs):
            _ops.mc_integrate(None, dim, N, domain, key, first, stats=stats[i], desc=d, out=out[i])
        self._last_stats = stats
        self._last_shape = fn.batch_shape
        if fn.batch_shape == (1,):
            _warn_scalar_deprecation()
            return torch.squeeze(out)
        return out.reshape(fn.batch_shape)

    def std_error(self):
        """Standard error of the last fused integral (one device read); an
        array of the batch shape for a parameter batch."""
        if self._last_stats is None:
            return None
        st = self._last_stats.double().reshape(-1, 3).cpu().numpy()
        s1, s2, n = st[:, 0], st[:, 1], st[:, 2]
        mean = s1 / n
        var = np.maximum(s2 / n - mean * mean, 0.0)
        err = np.sqrt(var / np.maximum(n - 1.0, 1.0))
        if self._last_stats.dim() == 1:
            self.last_std_error = float(err[0])
        else:
            self.last_std_error = err.reshape(self._last_shape)
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
        """Integrate function with dim and N fixed (monte_carlo.py:108-225).

        As in the reference, every call draws new sample points from one
        random stream started at ``seed``.  For a built-in integrand the fused
        step (sampling, integrand, reduction, finalize) is captured into a HIP
        graph on the first call per (integrand, domain dtype and device); the
        graph reads the domain and the stream position from device buffers,
        so a call copies the domain, stores the position and replays, with no
        other host work.  Other callables (and domains that require grad)
        run the eager path on the same stream.

        Returns:
            function(fn, integration_domain)
        """
        integration_domain = _setup_integration_domain(dim, integration_domain, backend)
        self._check_inputs(dim=dim, N=N, integration_domain=integration_domain)
        rng = RNG(backend="torch", seed=seed)
        graphs = {}

        def compiled_integrate(fn, integration_domain):
            dom = _float_domain(_setup_integration_domain(dim, integration_domain, backend))
            fused = (is_builtin(fn, dim) and not is_batched(fn) and dom.is_cuda
                     and not (torch.is_grad_enabled() and dom.requires_grad))
            if not fused:
                return self.integrate(fn, dim, N, dom, rng=rng, backend=backend)
            key = (bytes(fn.descriptor(dim)), dom.dtype, dom.device)
            entry = graphs.get(key)
            if entry is None:
                entry = self._capture_step(fn, dim, N, dom, rng.key)
                graphs[key] = entry
            graph, sdom, soff, sout, sstats, _ws = entry
            sdom.copy_(dom)
            soff.fill_(rng.reserve(N))
            graph.replay()
            self._nr_of_fevals = N
            self._last_stats = sstats.clone()
            self.last_std_error = None
            return sout.clone()

        return compiled_integrate

    @staticmethod
    def _capture_step(fn, dim, N, domain, key):
        """(graph, static domain, stream position, result, stats, workspace) of one built-in
        integrand's fused Monte Carlo step; the entry owns every buffer the graph uses."""
        dev = domain.device
        sdom = domain.detach().clone()
        soff = torch.zeros((), dtype=torch.int64, device=dev)
        sout = torch.empty((), dtype=sdom.dtype, device=dev)
        sstats = torch.empty(3, dtype=torch.float64, device=dev)
        ws = _ops.mc_workspace(dim, N, sdom)
        # eager warm-up (lazy initialisation outside the capture)
        _ops.mc_integrate(fn, dim, N, sdom, key, stats=sstats, offset_dev=soff, ws=ws, out=sout)
        torch.cuda.synchronize(dev)
        graph = torch.cuda.CUDAGraph()
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            with torch.cuda.graph(graph):
                _ops.mc_integrate(fn, dim, N, sdom, key, stats=sstats, offset_dev=soff, ws=ws, out=sout)
        return (graph, sdom, soff, sout, sstats, ws)
