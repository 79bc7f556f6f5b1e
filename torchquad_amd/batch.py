"""Batched form: many parameterised integrals in one launch per LDS class.

A :class:`Batch` takes a list of integrals -- rule, built-in integrand,
dimension, total point count N (rounded per axis exactly as the reference's
integrators round it) and domain -- plans them once on the host (C ABI
``tq_batch_plan``), keeps the plan on the device and runs all of them with
one wave per tile (one kernel per LDS class) plus one per-integral
reduction (``tq_batch_run``).  The launches have no host synchronisation and
no memset, so :meth:`Batch.graph` captures them in a HIP graph for replay.

The reference has no batched integrator; this mirrors its vectorised /
``Batch1DIntegrator`` use (base_integrator.py:77-89, tutorial.rst:800-888)
for independent integrals.
"""
import ctypes
import math

import numpy as np
import torch

from . import _lib, _ops
from ._lib import call, lib, ptr, stream_ptr
from .gl_nodes import fp32_pairs, gauss_legendre
from .integrands import BuiltinIntegrand

RULE_NAMES = {
    "trapezoid": _lib.RULE_TRAPEZOID,
    "simpson": _lib.RULE_SIMPSON,
    "boole": _lib.RULE_BOOLE,
    "gauss_legendre": _lib.RULE_GAUSS_LEGENDRE,
    "gausslegendre": _lib.RULE_GAUSS_LEGENDRE,
}

DEFAULT_TILE_POINTS = 8192


def n_per_axis(rule, dim, N):
    """Points per axis after the reference's N adjustment for `rule`.

    trapezoid / gauss_legendre: int(N^(1/dim) + 1e-8) (integration_grid.py:64);
    simpson: odd, >= 3 (simpson.py:54-81); boole: 1 + 4k, >= 5 (boole.py:56-84).
    """
    n = int(N ** (1.0 / dim) + 1e-8)
    if rule == _lib.RULE_SIMPSON:
        if n < 3:
            return 3
        if n % 2 != 1:
            return n - 1
    if rule == _lib.RULE_BOOLE:
        if n < 5:
            return 5
        if (n - 1) % 4 != 0:
            return n - ((n - 1) % 4)
    if n < 2 and rule != _lib.RULE_GAUSS_LEGENDRE:
        raise ValueError("N has to be > 1.")
    return max(n, 1)


_gl_cache = {}


def _gl_table(n, device):
    key = (int(n), str(device))
    t = _gl_cache.get(key)
    if t is None:
        t = gauss_legendre(int(n), device=device)
        _gl_cache[key] = t
    return t


class Batch:
    """A planned batch of integrals, resident on one device.

    Args:
        items: iterable of dicts with keys ``rule`` (name or code), ``fn``
            (a :class:`~torchquad_amd.integrands.BuiltinIntegrand`), ``dim``,
            ``domain`` ([[a, b]] * dim) and either ``N`` (total points, the
            integrators' argument) or ``n`` (points per axis).
        dtype: torch.float32 or torch.float64 (points and integrand).
        tile_points: points per tile (None: autotuned / default).
        gl_tables: callable (n, device) -> (nodes, weights), FP64 device
            tensors on [-1, 1], for the Gaussian items (default: the
            Gauss-Legendre generator); a custom Gaussian rule's tables.
    """

    def __init__(self, items, dtype=torch.float32, device=None, tile_points=None, gl_tables=None):
        if not torch.cuda.is_available():
            raise _lib.NativeUnavailableError("the batched form runs on a ROCm GPU only")
        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
        self.dtype = dtype
        self.dt = _lib.dtype_code(dtype)
        items = list(items)
        self.count = len(items)
        arr = (_lib.BatchItem * max(self.count, 1))()
        gl_ns = []
        self.evals = 0
        self.n = np.zeros(self.count, dtype=np.int64)
        self.dims = np.zeros(self.count, dtype=np.int64)
        self.families = np.zeros(self.count, dtype=np.int64)
        for i, it in enumerate(items):
            rule = it["rule"]
            rule = RULE_NAMES[rule.lower()] if isinstance(rule, str) else int(rule)
            fn = it["fn"]
            if not isinstance(fn, (BuiltinIntegrand, _OneSet)):
                raise TypeError("batched integrals need built-in integrands (torchquad_amd.integrands)")
            dim = int(it["dim"])
            n = int(it["n"]) if "n" in it else n_per_axis(rule, dim, int(it["N"]))
            dom = np.asarray(it.get("domain", [[-1.0, 1.0]] * dim), dtype=np.float64)
            if dom.shape != (dim, 2):
                raise ValueError(f"item {i}: domain shape {dom.shape} != ({dim}, 2)")
            if np.any(dom[:, 1] < dom[:, 0]):
                raise ValueError(f"item {i}: integration_domain has invalid boundary values")
            d = fn.descriptor(dim)
            e = arr[i]
            e.rule, e.family, e.dim, e.n = rule, d.family, dim, n
            e.gl_offset = 0
            for k in range(dim):
                e.domain[2 * k] = dom[k, 0]
                e.domain[2 * k + 1] = dom[k, 1]
            e.c0, e.c1, e.c2, e.c3 = d.c0, d.c1, d.c2, d.c3
            for k in range(_lib.TQ_MAX_DIM):
                e.a[k] = d.a[k]
                e.u[k] = d.u[k]
            if rule == _lib.RULE_GAUSS_LEGENDRE:
                gl_ns.append((i, n))
            self.evals += n**dim
            self.n[i], self.dims[i], self.families[i] = n, dim, d.family
        # Gauss-Legendre node/weight tables, one per distinct n, concatenated
        self.gl_x = self.gl_w = self.gl_p32 = None
        if gl_ns:
            offsets = {}
            parts_x, parts_w = [], []
            off = 0
            for _, n in gl_ns:
                if n not in offsets:
                    offsets[n] = off
                    x, w = (gl_tables or _gl_table)(n, self.device)
                    parts_x.append(x)
                    parts_w.append(w)
                    off += n
            self.gl_x = torch.cat(parts_x)
            self.gl_w = torch.cat(parts_w)
            if dtype == torch.float32:
                self.gl_p32 = fp32_pairs(self.gl_x, self.gl_w)
            for i, n in gl_ns:
                arr[i].gl_offset = offsets[n]
        self._items = arr
        self.tile_points = int(tile_points or DEFAULT_TILE_POINTS)
        self._plan(self.tile_points)
        self.out = torch.empty(self.count, dtype=dtype, device=self.device)
        self.out64 = torch.empty(self.count, dtype=torch.float64, device=self.device)

    def _plan(self, tile_points):
        desc_bytes = ctypes.c_int64()
        n_tiles = ctypes.c_int64()
        call("tq_batch_plan_sizes", ctypes.cast(self._items, ctypes.c_void_p), self.count, self.dt, int(tile_points),
             ctypes.byref(desc_bytes), ctypes.byref(n_tiles))
        desc = np.zeros(max(desc_bytes.value, 8), dtype=np.uint8)
        tiles = np.zeros(max(n_tiles.value, 1) * ctypes.sizeof(_lib.BatchTile), dtype=np.uint8)
        tb = np.zeros(max(self.count, 1), dtype=np.int64)
        tc = np.zeros(max(self.count, 1), dtype=np.int64)
        self.class_offsets = np.zeros(5, dtype=np.int64)
        call("tq_batch_plan", ctypes.cast(self._items, ctypes.c_void_p), self.count, self.dt, int(tile_points),
             desc.ctypes.data_as(ctypes.c_void_p), tiles.ctypes.data_as(ctypes.c_void_p),
             tb.ctypes.data_as(ctypes.c_void_p), tc.ctypes.data_as(ctypes.c_void_p),
             self.class_offsets.ctypes.data_as(ctypes.c_void_p))
        dev = self.device
        self.n_tiles = int(n_tiles.value)
        self.desc = torch.from_numpy(desc).to(dev)
        self.tiles = torch.from_numpy(tiles).to(dev)
        self.tile_begin = torch.from_numpy(tb).to(dev)
        self.tile_count = torch.from_numpy(tc).to(dev)
        self.tile_part = torch.empty(max(self.n_tiles, 1), dtype=torch.float64, device=dev)
        self.tile_points = int(tile_points)

    def retile(self, tile_points=None):
        """Re-plan with another tile size."""
        if tile_points is not None and int(tile_points) != self.tile_points:
            self._plan(int(tile_points))

    def run(self):
        """Launch the batch on torch's current stream; returns (out, out64)."""
        call("tq_batch_run", self.dt, ptr(self.desc), self.count, ptr(self.tiles), self.n_tiles,
             ptr(self.tile_begin), ptr(self.tile_count), ptr(self.gl_x), ptr(self.gl_w), ptr(self.gl_p32),
             self.gl_x.numel() if self.gl_x is not None else 0, ptr(self.tile_part), ptr(self.out), ptr(self.out64),
             self.class_offsets.ctypes.data_as(ctypes.c_void_p), stream_ptr(self.device))
        return self.out, self.out64

    def graph(self):
        """Capture run() in a HIP graph; returns the torch.cuda.CUDAGraph."""
        s = torch.cuda.Stream(device=self.device)
        s.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(s):
            self.run()  # warm-up outside capture
        torch.cuda.current_stream(self.device).wait_stream(s)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.run()
        return g


class ParamBatchRunner:
    """Integrals of a built-in family's parameter batch over one grid.

    Each of the prod(S) parameter sets of ``fn`` (``fn.batch_shape == S``)
    is one integral of a :class:`Batch` with the same rule, n, dimension and
    domain; plans are cached per (rule, n, parameters, domain, dtype,
    device), so a repeated call is one launch pair and no host work beyond
    it.  The reference's counterpart is a vectorised integrand returning
    ``[N, *S]`` (base_integrator.py:77-89).
    """

    def __init__(self, max_plans=16):
        self._plans = {}
        self._max = int(max_plans)

    def run(self, rule, fn, dim, n, domain, domain_host, gl_tables=None, tables_key=None):
        """[*S] integrals (domain dtype); domain_host: float64 numpy [dim, 2];
        tables_key: identifies what gl_tables returns (part of the plan key)."""
        key = (int(rule), int(n), fn.key(), np.ascontiguousarray(domain_host).tobytes(), domain.dtype,
               str(domain.device), tables_key)
        b = self._plans.get(key)
        if b is None:
            items = [{"rule": int(rule), "fn": f1, "dim": int(dim), "n": int(n), "domain": domain_host}
                     for f1 in _split(fn)]
            b = Batch(items, dtype=domain.dtype, device=domain.device, gl_tables=gl_tables)
            if len(self._plans) >= self._max:
                self._plans.pop(next(iter(self._plans)))
            self._plans[key] = b
        out, _ = b.run()
        return out.clone().reshape(fn.batch_shape)


class _OneSet:
    """One parameter set of a batched family, as a scalar descriptor source."""

    def __init__(self, desc):
        self._desc = desc

    def descriptor(self, dim=None):
        return self._desc


def _split(fn):
    return [_OneSet(d) for d in fn.descriptors(fn.dim)]


def integrate_batch(items, dtype=torch.float32, device=None, tile_points=None):
    """Plan and run a batch once; returns the float64 results tensor."""
    b = Batch(items, dtype=dtype, device=device, tile_points=tile_points)
    _, out64 = b.run()
    return out64
