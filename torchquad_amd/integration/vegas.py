"""VEGAS+ (reference: torchquad/integration/vegas.py:13-362).

The iteration structure, budget split, warmup, abort rules and the EQ 30-32
estimators follow the reference.  The work of an iteration runs on the GPU
with the map and stratification state resident there:

* built-in integrand: get_NH, sampling, map, integrand, histograms,
  per-cube sums and the map/stratification updates are device kernels with
  no host synchronisation; the host reads the per-iteration results once
  every five iterations, where the reference's abort check needs them;
* other callables: the device writes the iteration's points, ``fn`` runs
  on them, and the device consumes the values (one read of the sample count
  per iteration, as the reference's repeat_interleave also does).
  Gradients with respect to the domain and to tensors captured by ``fn``
  flow through the per-iteration results (the map and stratification
  statistics are detached, as in the reference).
"""
import math
import warnings

import torch

from .. import _vegas_dev
from .._lib import call, ptr
from ..integrands import is_builtin
from ..utils.set_log_level import logger
from .base_integrator import BaseIntegrator
from .rng import RNG
from .utils import _float_domain, _setup_integration_domain
from .vegas_map import VEGASMap, report_map_flags
from .vegas_stratification import VEGASStratification, rng_is_native


class VEGAS(BaseIntegrator):
    """VEGAS Enhanced (arXiv:2009.05112).  EQ <n> refers to that paper."""

    def __init__(self):
        super().__init__()

    def integrate(
        self,
        fn,
        dim,
        N=10000,
        integration_domain=None,
        seed=None,
        rng=None,
        use_grid_improve=True,
        eps_rel=0,
        eps_abs=0,
        max_iterations=20,
        use_warmup=True,
        backend=None,
        N_intervals=None,
        N_strat=None,
        warm_start=None,
    ):
        """Integrate fn with VEGAS+ (vegas.py:30-159).

        Args are the reference's, plus (all optional):
            N_intervals: map bins per axis; default the reference's
                max(2, N_increment // 10) (vegas.py:117).
            N_strat: stratification steps per axis; default the reference's
                min(1000, int((N_increment / 4)^(1/dim)))
                (vegas_stratification.py:27-28).
            warm_start: a ``state_dict()`` of an earlier run with the same
                dim, N_intervals and N_strat; the map edges and the
                stratification's dh start from it (typically with
                use_warmup=False).

        Returns:
            tensor: integral value
        """
        self._setup(fn, dim, N, integration_domain, seed, rng, use_grid_improve, eps_rel, eps_abs, max_iterations,
                    backend, N_intervals, N_strat, warm_start)
        for _ in self._steps(use_warmup):
            pass
        return self._finish()

    def _setup(self, fn, dim, N, integration_domain, seed, rng, use_grid_improve, eps_rel, eps_abs, max_iterations,
               backend, N_intervals, N_strat=None, warm_start=None):
        """Checks, budget split, RNG and resident map/stratification state (vegas.py:30-128)."""
        self._check_inputs(dim=dim, N=N, integration_domain=integration_domain)
        logger.debug("VEGAS integrating a %d-dimensional fn with %d points", dim, N)
        self._dim = dim
        self._nr_of_fevals = 0
        self._max_iterations = max_iterations
        self._eps_rel = eps_rel
        self._eps_abs = eps_abs
        self.use_grid_improve = use_grid_improve
        self.N = N
        self._starting_N = N // (self._max_iterations + 5)
        self._N_increment = N // (self._max_iterations + 5)
        self._fn = fn
        integration_domain = _setup_integration_domain(dim, integration_domain, backend)
        self.backend = "torch"
        domain = _float_domain(integration_domain)
        self.dtype = domain.dtype
        if rng is None:
            rng = RNG(backend="torch", seed=seed)
        elif seed is not None:
            raise ValueError("seed and rng cannot both be passed")
        self.rng = rng
        self._domain = domain
        self._grad = torch.is_grad_enabled() and domain.requires_grad
        self._fused = is_builtin(fn, dim) and rng_is_native(rng) and not self._grad

        if N_intervals is None:
            N_intervals = max(2, self._N_increment // 10)
        if N_strat is None:
            n_strat = _vegas_dev.n_strat_for(self._N_increment, dim)
        else:
            n_strat = int(N_strat)
            if n_strat < 1 or n_strat**dim > 1 << 31:
                raise ValueError("N_strat must be >= 1 with N_strat**dim < 2**31")
        # The device state of the previous call is reused when its shape, dtype, device and stream
        # match: one fill restores the freshly built state instead of about twenty allocations.
        key = (int(dim), int(N_intervals), max(n_strat, 1), self.dtype, domain.device,
               torch.cuda.current_stream(domain.device).cuda_stream if domain.is_cuda else None)
        cached = getattr(self, "_dev_cache", None)
        if cached is not None and cached[0] == key:
            dev = cached[1]
            dev.reset()
        else:
            dev = _vegas_dev.VegasDevice(dim, N_intervals, max(n_strat, 1), self.dtype, domain.device)
            self._dev_cache = (key, dev)
        self._dev = dev
        self.map = VEGASMap(N_intervals, dim, self.backend, self.dtype, _device=dev)
        self.strat = VEGASStratification(self._N_increment, dim, rng, self.backend, self.dtype, _device=dev)
        self._desc = fn.descriptor(dim) if self._fused else None
        self._flags_seen = [0, 0, 0, 0]
        if warm_start is not None:
            self._load_state(warm_start)

        self.results = []
        self.sigma2 = []
        self._slots = []
        self._prev_slot = -1
        self._slot_counter = 0
        self.it = 0

    def _steps(self, use_warmup):
        """Generator over the iterations (vegas.py:130-146).

        Yields after the warm-up and after each iteration's launches; the
        value is True when the next step will read the device (the abort
        check every fifth iteration), so a scheduler can wait for the
        device without blocking before resuming.
        """
        if use_warmup:
            self._warmup_grid(5, self._starting_N // 5)
            yield False
        while True:
            self.it = self.it + 1
            self._run_iteration()
            yield self.it % 5 == 0
            if self._check_abort_conditions():
                break

    def _finish(self):
        """Final statistics and dh normalisation; returns the result tensor (vegas.py:147-159)."""
        self._sync_results()
        call("tq_vegas_normalize_dh", self._dev.struct, self._prev_slot, self._dev.stream())
        result = self._get_result()
        if isinstance(result, torch.Tensor):
            return result.to(self.dtype)
        return torch.tensor(result, dtype=self.dtype, device=self._domain.device)

    # -- checkpoint ----------------------------------------------------------
    def state_dict(self):
        """Map edges and stratification dh after integrate(), for a warm start.

        The reference keeps this state in object attributes and re-creates it
        on every call (vegas.py:118-128); here it can be carried to a later
        integrate(..., warm_start=state) with the same dim, N_intervals and
        N_strat.
        """
        dev = self._dev
        return {
            "dim": dev.dim,
            "N_intervals": dev.n_intervals,
            "N_strat": dev.n_strat,
            "x_edges": dev.x_edges.detach().clone(),
            "dx_edges": dev.dx_edges.detach().clone(),
            "dh": dev.dh.detach().clone(),
        }

    def _load_state(self, state):
        dev = self._dev
        got = (int(state["dim"]), int(state["N_intervals"]), int(state["N_strat"]))
        want = (dev.dim, dev.n_intervals, dev.n_strat)
        if got != want:
            raise ValueError(f"warm_start has (dim, N_intervals, N_strat) = {got}, this run needs {want}")
        for name in ("x_edges", "dx_edges", "dh"):
            src = torch.as_tensor(state[name], dtype=torch.float64, device=dev.device)
            dst = getattr(dev, name)
            if tuple(src.shape) != tuple(dst.shape):
                raise ValueError(f"warm_start {name} has shape {tuple(src.shape)}, expected {tuple(dst.shape)}")
            dst.copy_(src)

    # -- device plumbing ---------------------------------------------------
    def _next_slot(self):
        s = self._slot_counter % _vegas_dev.MAX_ITERS
        self._slot_counter += 1
        return s

    def _check_flags(self):
        flags = self._dev.read_flags()
        if flags[3]:
            raise RuntimeError("VEGAS sample buffers overflowed; this indicates an invalid stratification state")
        report_map_flags(self._flags_seen, flags)
        self._flags_seen = flags

    def _user_iteration(self, slot, warmup, n_samples):
        """points -> fn -> consume; returns (differentiable result or None)."""
        dev = self._dev
        if warmup:
            total = n_samples
        else:
            total = int(dev.offsets[-1].item())
        dom = self._domain
        unit = torch.tensor([[0.0, 1.0]] * self._dim, dtype=dom.dtype, device=dom.device)
        x = torch.empty((total, self._dim), dtype=dom.dtype, device=dom.device)
        jac = torch.empty((total,), dtype=torch.float64, device=dom.device)
        ids = torch.empty((total, self._dim), dtype=torch.int32, device=dom.device)
        if rng_is_native(self.rng):
            first = self.rng.reserve(total if warmup else dev.capacity_for(self._starting_N))
            u = None
            seed = self.rng.key
        else:
            u = self.rng.uniform(size=[total, self._dim], dtype=dom.dtype).to(dom.device).contiguous()
            first, seed = 0, 0
        call("tq_vegas_points", dev.struct, ptr(unit if self._grad else dom.detach()), seed, first, 1 if warmup else 0,
             n_samples if warmup else 0, ptr(u), ptr(x), ptr(jac), ptr(ids), None, dev.stream())
        starts = dom[:, 0]
        sizes = dom[:, 1] - starts
        if self._grad:
            x = x * sizes + starts
        values, num = self.evaluate_integrand(self._fn, x)
        self._nr_of_fevals += num
        values = values.squeeze()
        if values.dim() != 1:
            raise ValueError("VEGAS supports only scalar integrands")
        vals = values.detach().to(dom.dtype).contiguous()
        call("tq_vegas_consume", dev.struct, ptr(vals), ptr(unit if self._grad else dom.detach()), ptr(jac),
             ptr(ids), slot, 1 if self.use_grid_improve else 0, 1 if warmup else 0, n_samples if warmup else 0,
             dev.stream())
        if warmup or not (self._grad or values.requires_grad):
            return None
        # differentiable iteration result: sum_s V/nh(s) * jac_s * vol * f_s
        volume = torch.prod(sizes)
        nh = dev.nh
        per_sample = torch.repeat_interleave(self.strat.V_cubes / nh.to(torch.float64), nh, output_size=total)
        coef = (per_sample * jac).to(values.dtype)
        return torch.sum(coef * values) * volume

    def _warmup_grid(self, warmup_N_it=5, N_samples=1000):
        """Adapt the map on plain uniform samples; results discarded (vegas.py:211-266)."""
        self.alpha = 0.5
        dev = self._dev
        dev.ensure_capacity(N_samples)
        for _ in range(warmup_N_it):
            slot = self._next_slot()
            if self._fused:
                first = self.rng.reserve(N_samples)
                call("tq_vegas_iterate", dev.struct, self._desc, ptr(self._domain), self.rng.key, first, slot, 1, 1,
                     N_samples, dev.stream())
                self._nr_of_fevals += N_samples
            else:
                self._user_iteration(slot, True, N_samples)
        self.results.clear()
        self.sigma2.clear()

    def _run_iteration(self):
        """One VEGAS iteration (vegas.py:268-315); results stay on the device."""
        dev = self._dev
        slot = self._next_slot()
        dev.ensure_capacity(dev.capacity_for(self._starting_N))
        call("tq_vegas_nh", dev.struct, int(self._starting_N), self._prev_slot, slot, None, dev.stream())
        diff = None
        if self._fused:
            first = self.rng.reserve(dev.capacity_for(self._starting_N))
            call("tq_vegas_iterate", dev.struct, self._desc, ptr(self._domain), self.rng.key, first, slot,
                 1 if self.use_grid_improve else 0, 0, 0, dev.stream())
        else:
            diff = self._user_iteration(slot, False, 0)
        self._prev_slot = slot
        self._slots.append((slot, diff))
        self.results.append(None)
        self.sigma2.append(None)

    def _sync_results(self):
        """Fill results/sigma2 of the pending iterations from the device."""
        pending = [i for i, r in enumerate(self.results) if r is None]
        if not pending:
            return
        stats = self._dev.iter_stats[[self._slots[i][0] for i in pending]].tolist()
        # with domain gradients the device statistics are taken on the unit cube
        scale = 1.0
        if self._grad:
            scale = float(torch.prod(self._domain[:, 1] - self._domain[:, 0]).detach())
        for i, st in zip(pending, stats):
            diff = self._slots[i][1]
            self.results[i] = diff if diff is not None else st[0] * scale
            self.sigma2[i] = abs(st[1]) * scale * scale
            if self._fused:
                self._nr_of_fevals += int(st[2])
        self._check_flags()

    def _check_abort_conditions(self):
        """Stop or adjust the evaluations per iteration every fifth iteration (vegas.py:161-209)."""
        if self.it % 5 > 0:
            return False
        self._sync_results()
        res_abs = abs(float(self._get_result()))
        err = float(self._get_error())
        chi2 = float(self._get_chisq())
        if (err <= self._eps_rel * res_abs or err <= self._eps_abs) and chi2 / 5.0 < 1.0:
            return True
        if chi2 / 5.0 < 1.0:
            if res_abs == 0.0:
                self._starting_N += self._N_increment
            else:
                acc = err / res_abs
                self._starting_N = min(
                    self._starting_N + self._N_increment,
                    int(self._starting_N * math.sqrt(acc / (self._eps_rel + 1e-8))),
                )
        elif chi2 / 5.0 > 1.0:
            self._starting_N += self._N_increment
        if self._nr_of_fevals + self._starting_N * 5 > self.N:
            return True
        if self.it + 5 > self._max_iterations:
            return True
        self.results = []
        self.sigma2 = []
        self._slots = []
        return False

    def _get_result(self):
        """Inverse-variance weighted mean, EQ 30 (vegas.py:318-335)."""
        self._sync_results()
        if any(sig2 == 0.0 for sig2 in self.sigma2):
            return sum(self.results) / len(self.results)
        res_num = sum(res / sig2 for res, sig2 in zip(self.results, self.sigma2))
        res_den = sum(1.0 / sig2 for sig2 in self.sigma2)
        return res_num / res_den

    def _get_error(self):
        """Error estimate, EQ 31 (vegas.py:337-346)."""
        res = sum(1.0 / sig2 for sig2 in self.sigma2 if sig2 != 0.0)
        return self.sigma2[0] if res == 0 else 1.0 / math.sqrt(res)

    def _get_chisq(self):
        """Chi square, EQ 32 (vegas.py:348-362).  The reference divides backend
        tensors, so an entry with zero variance that differs from the result
        contributes inf."""
        I_final = float(self._get_result())
        return sum(
            (((float(res) - I_final) ** 2 / sig2 if sig2 != 0.0 else math.inf)
             for res, sig2 in zip(self.results, self.sigma2) if float(res) != I_final),
            0.0,
        )
