"""Serving form: a continuous-batching request loop on one GPU.

Requests for independent integrals -- tensor-product rules over built-in
integrands of any dim / N / domain, and VEGAS integrals -- arrive at any
time (:meth:`IntegrationServer.submit_grid`, :meth:`submit_vegas`, from
any thread).  Each scheduling round (:meth:`IntegrationServer.step`):

1. polls the device events of the work in flight and hands out finished
   results (no blocking synchronisation);
2. admits waiting requests in arrival order while their device-memory
   reservation fits the HBM budget; a request larger than the whole budget
   is rejected with :class:`AdmissionError`;
3. packs every admitted grid request into one batch -- host plan
   (``tq_batch_plan``), asynchronous upload from pinned buffers, the
   tile kernels plus the per-integral reduction (``tq_batch_run``),
   asynchronous download of the results -- on one of ``max_inflight``
   streams, so a batch is planned and launched while earlier ones still
   run.  Gauss-Legendre tables live at fixed offsets of one server-wide
   arena, generated once per n on the table stream; a request waits in the
   admitted set until the host has seen its tables' event, so batches carry
   no per-table stream dependencies;
4. advances each resident VEGAS request by one iteration (all VEGAS
   requests share one stream; with the two batch streams and the table
   stream the server stays within the device's four hardware queues).  Map
   and stratification state stay on the device between iterations; the
   every-fifth-iteration abort check runs only once the request's event
   shows the device has finished those iterations.

Requests come one at a time (:meth:`submit_grid`) or as packed arrays
(:meth:`submit_many`: numpy records in the layout of ``tq_batch_item``, so a
set of integrals costs one admission, one host plan and one launch, with no
per-integral Python work).

A grid integral's value depends only on the integral and the server's
``tile_points`` (its tiles and the fixed-order sum over them do not depend
on the other integrals of the batch); a VEGAS integral's value depends only
on its own arguments and seed.  So every result is the same alone and in
any mix of other requests.

Latency: the loop runs in Python, so a full pass of the cyclic garbage
collector over a large heap stalls it (tens of ms in a process holding
hundreds of thousands of objects).  A serving process should call
``gc.collect(); gc.freeze()`` once its long-lived state is built.

The reference has no serving layer; the integrals it serves are those of
Trapezoid/Simpson/Boole/GaussLegendre.integrate (trapezoid.py:12,
simpson.py:14, boole.py:14, gaussian.py:29) and VEGAS.integrate
(vegas.py:30-44).
"""
import ctypes
import itertools
import struct
import threading
import time
from collections import deque

import numpy as np
import torch

from . import _lib
from ._lib import NativeUnavailableError, call, lib, ptr
from .batch import RULE_NAMES, n_per_axis
from .integrands import BuiltinIntegrand

ITEM_BYTES = ctypes.sizeof(_lib.BatchItem)
TILE_BYTES = ctypes.sizeof(_lib.BatchTile)

# tq_batch_item as a numpy record: requests are packed, concatenated and
# uploaded as arrays, so a batch costs no per-request Python work.
ITEM_DTYPE = np.dtype([
    ("rule", np.int32), ("family", np.int32), ("dim", np.int32), ("n", np.int32), ("gl_offset", np.int64),
    ("domain", np.float64, (2 * _lib.TQ_MAX_DIM,)), ("c0", np.float64), ("c1", np.float64), ("c2", np.float64),
    ("c3", np.float64), ("a", np.float64, (_lib.TQ_MAX_DIM,)), ("u", np.float64, (_lib.TQ_MAX_DIM,)),
])
assert ITEM_DTYPE.itemsize == ITEM_BYTES
_ITEM_HEAD = struct.Struct("<iiiiq")
_ZERO_DOMAIN = bytes(16 * _lib.TQ_MAX_DIM)
_INTEGRAND_PARAMS = _lib.Integrand.c0.offset  # c0..c3, a, u: the tail of both structs
assert ctypes.sizeof(_lib.Integrand) - _INTEGRAND_PARAMS == ITEM_BYTES - ITEM_DTYPE.fields["c0"][1]


def n_per_axis_array(rule, dim, N):
    """Vectorised n_per_axis (the reference's N -> n rounding per rule)."""
    rule, dim, N = (np.asarray(v, dtype=np.int64) for v in (rule, dim, N))
    n = np.floor(N.astype(np.float64) ** (1.0 / dim) + 1e-8).astype(np.int64)
    simpson = rule == _lib.RULE_SIMPSON
    n = np.where(simpson & (n < 3), 3, np.where(simpson & (n % 2 != 1), n - 1, n))
    boole = rule == _lib.RULE_BOOLE
    n = np.where(boole & (n < 5), 5, np.where(boole & ((n - 1) % 4 != 0), n - ((n - 1) % 4), n))
    if np.any((n < 2) & (rule != _lib.RULE_GAUSS_LEGENDRE)):
        raise ValueError("N has to be > 1.")
    return np.maximum(n, 1)


class AdmissionError(RuntimeError):
    """The request can never fit the server's HBM budget."""


class CancelledError(RuntimeError):
    """The request was cancelled before its result was delivered."""


class Ticket:
    """Handle of one request: wait for its result, or cancel it.  A ticket of
    ``submit_many`` covers ``count`` integrals and its result is an array."""

    _ids = itertools.count()

    def __init__(self, kind, nbytes, evals, count=1):
        self.id = next(Ticket._ids)
        self.count = int(count)
        self.kind = kind
        self.nbytes = int(nbytes)
        self.evals = int(evals)
        self.t_submit = time.perf_counter()
        self.t_arrival = self.t_submit
        self.t_admit = None
        self.t_done = None
        self.value = None
        self.error = None
        self.cancelled = False
        self.state = None  # VEGAS requests: the run's state_dict(), for a later warm start
        self.rule = None  # grid requests: rule code
        self._done = False
        self._event = None  # made by the first waiter (most tickets are read once done)

    _lock = threading.Lock()

    def done(self):
        return self._done

    @property
    def latency(self):
        """Seconds from arrival to delivery (None until done)."""
        return None if self.t_done is None else self.t_done - self.t_arrival

    def result(self, timeout=None):
        """The integral (float64 Python float); raises the request's error."""
        with Ticket._lock:
            if not self._done and self._event is None:
                self._event = threading.Event()
            ev = self._event
        if not self._done and not ev.wait(timeout):
            raise TimeoutError(f"request {self.id} not finished within {timeout} s")
        if self.cancelled:
            raise CancelledError(f"request {self.id} was cancelled")
        if self.error is not None:
            raise self.error
        return self.value

    def _finish(self, value=None, error=None):
        self.value = value
        self.error = error
        self.t_done = time.perf_counter()
        with Ticket._lock:
            self._done = True
            ev = self._event
        if ev is not None:
            ev.set()


class _GridRequest:
    """One ticket's integrals: packed items and their Gauss-Legendre n (0 if none)."""

    __slots__ = ("ticket", "items", "gl_n", "gl_unique", "gl_seq", "single")

    def __init__(self, ticket, items, gl_n, single):
        self.ticket = ticket
        self.items = items
        self.gl_n = gl_n
        self.gl_unique = np.unique(gl_n[gl_n > 0]) if len(gl_n) > 1 else gl_n[gl_n > 0]
        self.gl_seq = None  # table generation its tables wait for, once placed in the arena
        self.single = single


class _GlArena:
    """Gauss-Legendre tables of every n the server has used, at fixed offsets
    of one device buffer (float64 nodes and weights, plus FP32 (node, weight)
    pairs for a float32 server).

    A table is generated once, in place, on the arena's stream
    (``tq_gauss_legendre``); each round's new tables are followed by one event,
    and a request is launched only once the host has seen the event of its
    tables fire, so batches read the arena with no stream dependency.  Tables
    are never moved while batches may read them: the arena grows into new
    buffers (the old ones stay referenced by the batches in flight) and is
    cleared only while no batch is in flight.
    """

    def __init__(self, device, fp32, limit_bytes):
        self.device = device
        self.fp32 = fp32
        self.per = 24 if fp32 else 16
        self.limit = max(int(limit_bytes) // self.per, 1)
        self.cap = 0
        self.x = self.w = self.p = None
        self.offset = {}
        self.used = 0
        self.seq = 0  # last generation issued
        self.ready_seq = 0  # last generation seen complete
        self.floor_seq = 0  # generation every table needs (the last regrowth copy)
        self.pending = deque()
        self.stream = torch.cuda.Stream(device=device)

    def poll(self):
        while self.pending and self.pending[0][0].query():
            self.ready_seq = self.pending.popleft()[1]

    def idle(self):
        self.poll()
        return not self.pending

    def clear(self):
        """Forget every table (only while no batch reads the arena and no
        generation is in flight)."""
        self.offset.clear()
        self.used = 0

    def _grow(self, need):
        """Move to buffers that hold need more entries; the copy of the
        resident tables is ordered before the next generation's event."""
        cap = min(max(2 * self.cap, self.used + need, 1 << 16), max(self.limit, self.used + need))
        x = torch.empty(cap, dtype=torch.float64, device=self.device)
        w = torch.empty(cap, dtype=torch.float64, device=self.device)
        p = torch.empty(2 * cap, dtype=torch.float32, device=self.device) if self.fp32 else None
        self.stream.wait_stream(torch.cuda.current_stream(self.device))  # the allocations
        if self.used:
            with torch.cuda.stream(self.stream):
                for old, new, k in ((self.x, x, 1), (self.w, w, 1), (self.p, p, 2)):
                    if new is not None:
                        new[: k * self.used].copy_(old[: k * self.used])
                        old.record_stream(self.stream)
            self.floor_seq = self.seq + 1
        self.x, self.w, self.p, self.cap = x, w, p, cap

    def place(self, ns):
        """Start the tables of ns (distinct, > 0) that are not resident; returns
        the generation the caller must see complete, or None if the tables
        do not fit before the arena is cleared."""
        new = [int(n) for n in ns if int(n) not in self.offset]
        if new:
            need = sum(new)
            if self.used + need > self.cap:
                if self.used + need > self.limit and self.used:
                    return None
                self._grow(need)
            self.seq += 1
            with torch.cuda.stream(self.stream):
                for n in new:
                    o = self.used
                    call("tq_gauss_legendre", n, ptr(self.x[o:]), ptr(self.w[o:]), self.stream.cuda_stream)
                    if self.p is not None:
                        pairs = self.p[2 * o: 2 * (o + n)].view(n, 2)
                        pairs[:, 0].copy_(self.x[o: o + n])
                        pairs[:, 1].copy_(self.w[o: o + n])
                    self.offset[n] = (o, self.seq)
                    self.used += n
                ev = torch.cuda.Event()
                ev.record(self.stream)
            self.pending.append((ev, self.seq))
        return max([self.floor_seq] + [self.offset[int(n)][1] for n in ns])

    def offsets(self, gl_n):
        """Arena offsets of the tables of gl_n (entries > 0)."""
        u, inv = np.unique(gl_n, return_inverse=True)
        offs = np.fromiter((self.offset[int(n)][0] for n in u), dtype=np.int64, count=len(u))
        return offs[inv]


class _VegasRequest:
    __slots__ = ("ticket", "args", "integrator", "steps", "stream", "event", "wait")

    def __init__(self, ticket, args):
        self.ticket = ticket
        self.args = args
        self.integrator = None
        self.steps = None
        self.stream = None
        self.event = None
        self.wait = False


def _aligned(nbytes, to=256):
    return (int(nbytes) + to - 1) // to * to


class _Pinned:
    """A growable pinned host buffer and its device twin (bytes)."""

    def __init__(self, device):
        self.device = device
        self.host = torch.empty(0, dtype=torch.uint8)
        self.dev = torch.empty(0, dtype=torch.uint8, device=device)

    def reserve(self, nbytes):
        nbytes = max(int(nbytes), 8)
        if self.host.numel() < nbytes:
            size = 1 << (nbytes - 1).bit_length()
            self.host = torch.empty(size, dtype=torch.uint8, pin_memory=True)
            self.dev = torch.empty(size, dtype=torch.uint8, device=self.device)
        return self.host, self.dev


class _BatchSlot:
    """Buffers of one in-flight batch; reused once its event has fired."""

    def __init__(self, device, dtype):
        self.device = device
        self.dtype = dtype
        self.stream = torch.cuda.Stream(device=device)
        self.stage = _Pinned(device)  # plan outputs: descriptors | tiles | tile_begin | tile_count
        self.part = torch.empty(0, dtype=torch.float64, device=device)
        self.class_offsets = np.zeros(5, dtype=np.int64)
        self.out = torch.empty(0, dtype=dtype, device=device)
        self.out64 = torch.empty(0, dtype=torch.float64, device=device)
        self.host_out = torch.empty(0, dtype=torch.float64, pin_memory=True)
        self.event = None
        self.requests = None
        self.gl = None

    def busy(self):
        return self.requests is not None

    def ensure_outputs(self, count, n_tiles):
        if self.out.numel() < count:
            size = 1 << max(count - 1, 1).bit_length()
            self.out = torch.empty(size, dtype=self.dtype, device=self.device)
            self.out64 = torch.empty(size, dtype=torch.float64, device=self.device)
            self.host_out = torch.empty(size, dtype=torch.float64, pin_memory=True)
        if self.part.numel() < n_tiles:
            self.part = torch.empty(1 << max(n_tiles - 1, 1).bit_length(), dtype=torch.float64, device=self.device)


class IntegrationServer:
    """Continuous-batching integration server on one GPU.

    Args:
        device: the GPU (default: current).
        dtype: torch.float32 or torch.float64, the compute dtype of grid
            requests (VEGAS requests use the dtype of their domain).
        hbm_budget: bytes of device memory that admitted requests may
            reserve at once (default: half of the free memory).
        tile_points: points per tile of the batched kernel; fixed per
            server so that results do not depend on the mix.
        max_inflight: grid batches in flight at once (one stream each).
        max_vegas: VEGAS requests resident at once.
        max_batch: grid requests per batch.
        gl_cache_limit: bytes the Gauss-Legendre table arena may grow to
            (tables of every n used so far, 24 B per node for a float32
            server; cleared when full while nothing reads it).  The default
            4 GiB holds about 1.8e8 nodes: a mix of one-axis integrals up
            to n = 1e6 keeps every table resident, as the integrators'
            own per-n caches do.
    """

    def __init__(self, device=None, dtype=torch.float32, hbm_budget=None, tile_points=8192, max_inflight=2,
                 max_vegas=8, max_batch=8192, gl_cache_limit=4 << 30):
        if not torch.cuda.is_available():
            raise NativeUnavailableError("the serving loop runs on a ROCm GPU only")
        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
        self.dtype = dtype
        self.dt = _lib.dtype_code(dtype)
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            hbm_budget = free // 2
        self.hbm_budget = int(hbm_budget)
        self.reserved = 0
        self.peak_reserved = 0
        self.tile_points = int(tile_points)
        self.max_vegas = int(max_vegas)
        self.max_batch = int(max_batch)
        self._lock = threading.Lock()
        self._waiting = deque()
        self._admitted_grid = []
        self._vegas = []
        self._slots = [_BatchSlot(self.device, dtype) for _ in range(int(max_inflight))]
        self._next_slot = 0
        self.gl_cache_limit = int(gl_cache_limit)
        self._arena = _GlArena(self.device, self.dt == _lib.TQ_F32, self.gl_cache_limit)
        self._vegas_shared_stream = None
        self._thread = None
        self._stop = threading.Event()
        self.completed = []
        self.batches_launched = 0

    # -- submission ---------------------------------------------------------
    def submit_grid(self, rule, fn, dim, N=None, n=None, domain=None, arrival=None):
        """Queue a tensor-product integral of a built-in integrand; returns a Ticket.

        rule: "trapezoid" | "simpson" | "boole" | "gauss_legendre"; N is
        rounded per axis like the integrators do, or give n per axis.
        """
        code = RULE_NAMES[rule.lower()] if isinstance(rule, str) else int(rule)
        if not isinstance(fn, BuiltinIntegrand):
            raise TypeError("the server integrates built-in integrands (torchquad_amd.integrands)")
        dim = int(dim)
        if n is None:
            if N is None:
                raise ValueError("give N (total points) or n (points per axis)")
            n = n_per_axis(code, dim, int(N))
        dom = np.asarray(domain if domain is not None else [[-1.0, 1.0]] * dim, dtype=np.float64)
        if dom.shape != (dim, 2):
            raise ValueError(f"domain shape {dom.shape} != ({dim}, 2)")
        if np.any(dom[:, 1] < dom[:, 0]):
            raise ValueError("integration_domain has invalid boundary values")
        d = fn.descriptor(dim)
        # the record's bytes: header, zero-padded domain, then c0..c3, a, u,
        # which tq_integrand holds in the same layout after family and dim
        items = np.frombuffer(_ITEM_HEAD.pack(code, d.family, dim, int(n), 0) + dom.tobytes()
                              + _ZERO_DOMAIN[16 * dim:] + bytes(d)[_INTEGRAND_PARAMS:], dtype=ITEM_DTYPE)
        gl_n = np.array([int(n) if code == _lib.RULE_GAUSS_LEGENDRE else 0], dtype=np.int64)
        t = Ticket("grid", self._grid_bytes(items, gl_n), int(n) ** dim)
        t.rule = code
        if arrival is not None:
            t.t_arrival = arrival
        self._enqueue(_GridRequest(t, items, gl_n, True))
        return t

    def submit_many(self, rule, family, dim, domain, a, u, c=None, n=None, N=None, arrival=None):
        """Queue many tensor-product integrals at once; returns one Ticket whose
        result() is the float64 array of their values.

        Arrays of length R (scalars broadcast): rule codes, family codes
        (torchquad_amd._lib.FAM_*), dim, and n (points per axis) or N (total
        points, rounded per rule like the integrators); domain [R, dim_max, 2],
        a and u [R, k <= 16] and c [R, 4] (c0..c3) are the built-in family
        parameters as in tq_integrand.  Items are packed with numpy and
        admitted, planned and launched as one set.
        """
        rule = np.atleast_1d(np.asarray(rule, dtype=np.int64))
        # R from every argument that may carry it (per-integral scalars, or
        # the leading axis of domain / a / u / c)
        lead = [rule.shape, np.shape(np.atleast_1d(family)), np.shape(np.atleast_1d(dim)),
                np.shape(domain)[:-2], np.shape(a)[:-1], np.shape(u)[:-1]]
        lead += [np.shape(np.atleast_1d(v)) for v in (n, N) if v is not None]
        if c is not None:
            lead.append(np.shape(c)[:-1])
        R = int(np.broadcast_shapes(*[sh if len(sh) else (1,) for sh in lead])[0])
        rule = np.broadcast_to(rule, (R,))
        family = np.broadcast_to(np.asarray(family, dtype=np.int64), (R,))
        dim = np.broadcast_to(np.asarray(dim, dtype=np.int64), (R,))
        if np.any((dim < 1) | (dim > _lib.TQ_MAX_DIM)):
            raise ValueError(f"dim must be in 1..{_lib.TQ_MAX_DIM}")
        if np.any((family < 0) | (family >= 7)) or np.any((rule < 0) | (rule > 3)):
            raise ValueError("unknown rule or family code")
        if n is None:
            if N is None:
                raise ValueError("give N (total points) or n (points per axis)")
            n = n_per_axis_array(rule, dim, np.broadcast_to(np.asarray(N, dtype=np.int64), (R,)))
        n = np.broadcast_to(np.asarray(n, dtype=np.int64), (R,))
        dom = np.asarray(domain, dtype=np.float64)
        dom = np.broadcast_to(dom, (R,) + dom.shape[-2:])
        if dom.shape[1] > _lib.TQ_MAX_DIM or dom.shape[2] != 2:
            raise ValueError("domain must be [R, <=16, 2]")
        axis_ok = np.arange(dom.shape[1])[None, :] < dim[:, None]
        if np.any(axis_ok & (dom[:, :, 1] < dom[:, :, 0])):
            raise ValueError("integration_domain has invalid boundary values")
        items = np.zeros(R, dtype=ITEM_DTYPE)
        items["rule"], items["family"], items["dim"], items["n"] = rule, family, dim, n
        items["domain"][:, : 2 * dom.shape[1]] = np.where(axis_ok[:, :, None], dom, 0.0).reshape(R, -1)
        cc = np.broadcast_to(np.asarray(c if c is not None else [1.0, 0.0, 0.0, 0.0], dtype=np.float64), (R, 4))
        items["c0"], items["c1"], items["c2"], items["c3"] = cc[:, 0], cc[:, 1], cc[:, 2], cc[:, 3]
        for name, v in (("a", a), ("u", u)):
            v = np.asarray(v, dtype=np.float64)
            v = np.broadcast_to(v, (R,) + v.shape[-1:])
            items[name][:, : v.shape[1]] = v
        gl_n = np.where(rule == _lib.RULE_GAUSS_LEGENDRE, n, 0).astype(np.int64)
        evals = int(np.sum(n.astype(np.float64) ** dim))
        t = Ticket("grid", self._grid_bytes(items, gl_n), evals, count=R)
        if arrival is not None:
            t.t_arrival = arrival
        self._enqueue(_GridRequest(t, items, gl_n, False))
        return t

    def _grid_bytes(self, items, gl_n):
        """Device memory a set of grid items reserves: descriptors, tiles,
        partials and outputs.  (Gauss-Legendre tables live in the server's
        arena, at most gl_cache_limit bytes beside the budget.)"""
        desc_bytes = ctypes.c_int64()
        n_tiles = ctypes.c_int64()
        call("tq_batch_plan_sizes", ctypes.c_void_p(items.ctypes.data), len(items), self.dt, self.tile_points,
             ctypes.byref(desc_bytes), ctypes.byref(n_tiles))
        return int(desc_bytes.value + n_tiles.value * (TILE_BYTES + 8) + len(items) * (32 + ITEM_BYTES))

    def submit_vegas(self, fn, dim, N, domain=None, seed=0, N_intervals=None, max_iterations=20, eps_rel=0,
                     eps_abs=0, use_grid_improve=True, use_warmup=True, arrival=None, N_strat=None,
                     warm_start=None):
        """Queue a VEGAS integral of a built-in integrand; returns a Ticket.

        N_strat and warm_start are those of VEGAS.integrate; a warm start is
        the state_dict() of an earlier VEGAS run (e.g. a previous request's
        integrator, kept by the caller).
        """
        from ._vegas_dev import SB, n_strat_for

        if not isinstance(fn, BuiltinIntegrand):
            raise TypeError("the server integrates built-in integrands (torchquad_amd.integrands)")
        dim, N = int(dim), int(N)
        n_inc = N // (max_iterations + 5)
        ni = N_intervals if N_intervals is not None else max(2, n_inc // 10)
        nc = max(int(N_strat) if N_strat is not None else n_strat_for(n_inc, dim), 1) ** dim
        cap = int(1.25 * (N // 5 + 3 * nc + SB)) + SB
        blocks = cap // SB + 4
        nbytes = (8 * (dim * (ni + 1) + 3 * dim * ni) + nc * 48 + 4096 * 32 + cap * 8
                  + blocks * dim * ni * 12 + max(blocks, nc // 256 + 1) * 32)
        t = Ticket("vegas", nbytes, N)
        if arrival is not None:
            t.t_arrival = arrival
        args = dict(fn=fn, dim=dim, N=N, domain=domain, seed=seed, N_intervals=ni, max_iterations=max_iterations,
                    eps_rel=eps_rel, eps_abs=eps_abs, use_grid_improve=use_grid_improve, use_warmup=use_warmup,
                    N_strat=N_strat, warm_start=warm_start)
        self._enqueue(_VegasRequest(t, args))
        return t

    def _enqueue(self, req):
        if req.ticket.nbytes > self.hbm_budget:
            req.ticket._finish(error=AdmissionError(
                f"request needs {req.ticket.nbytes} B of device memory, more than the budget of {self.hbm_budget} B"))
            return
        with self._lock:
            self._waiting.append(req)

    def cancel(self, ticket):
        """Cancel a request: dropped if still waiting, result discarded if on the device."""
        ticket.cancelled = True
        with self._lock:
            for r in list(self._waiting):
                if r.ticket is ticket:
                    self._waiting.remove(r)
                    ticket._finish()
                    return True
        return False

    # -- scheduling ---------------------------------------------------------
    def pending(self):
        with self._lock:
            waiting = len(self._waiting)
        return waiting + len(self._admitted_grid) + len(self._vegas) + sum(s.busy() for s in self._slots)

    def step(self):
        """One scheduling round; returns the number of requests finished in it."""
        finished = self._poll_batches()
        self._admit()
        self._launch_grid()
        finished += self._advance_vegas()
        return finished

    def _release(self, ticket):
        self.reserved -= ticket.nbytes

    def _poll_batches(self):
        finished = 0
        for slot in self._slots:
            if slot.busy() and slot.event.query():
                total = sum(len(r.items) for r in slot.requests)
                vals = slot.host_out[:total].numpy()
                pos = 0
                for r in slot.requests:
                    k = len(r.items)
                    self._release(r.ticket)
                    if not r.ticket.cancelled:
                        r.ticket._finish(value=float(vals[pos]) if r.single else vals[pos:pos + k].copy())
                    else:
                        r.ticket._finish()
                    pos += k
                    self.completed.append(r.ticket)
                finished += len(slot.requests)
                slot.requests = None
                slot.gl = None
                slot.event = None
        return finished

    def _admit(self):
        admitted = sum(len(g.items) for g in self._admitted_grid)
        with self._lock:
            while self._waiting:
                r = self._waiting[0]
                if self.reserved + r.ticket.nbytes > self.hbm_budget:
                    break
                if isinstance(r, _VegasRequest) and len(self._vegas) >= self.max_vegas:
                    break
                if isinstance(r, _GridRequest) and admitted and admitted + len(r.items) > self.max_batch:
                    break
                self._waiting.popleft()
                self.reserved += r.ticket.nbytes
                self.peak_reserved = max(self.peak_reserved, self.reserved)
                r.ticket.t_admit = time.perf_counter()
                if isinstance(r, _GridRequest):
                    self._admitted_grid.append(r)
                    admitted += len(r.items)
                else:
                    self._vegas.append(r)

    def _launch_grid(self):
        if not self._admitted_grid:
            return
        arena = self._arena
        arena.poll()
        idle = not any(s.busy() for s in self._slots)
        for r in self._admitted_grid:  # start tables even while every slot is busy
            if r.gl_seq is None and len(r.gl_unique):
                r.gl_seq = arena.place(r.gl_unique)
                # Over the arena's limit: start afresh, but only while nothing
                # reads it (no batch in flight, no table being generated) and
                # no admitted request holds a placement that it has not
                # launched yet -- clearing those let two requests that do not
                # fit together take the arena from each other forever.
                if (r.gl_seq is None and idle and arena.idle()
                        and all(q.gl_seq is None for q in self._admitted_grid)):
                    arena.clear()
                    r.gl_seq = arena.place(r.gl_unique)
        slot = None
        for k in range(len(self._slots)):
            s = self._slots[(self._next_slot + k) % len(self._slots)]
            if not s.busy():
                slot = s
                self._next_slot = (self._next_slot + k + 1) % len(self._slots)
                break
        if slot is None:
            return
        reqs, later = [], []
        for r in self._admitted_grid:
            if len(r.gl_unique) and (r.gl_seq is None or max(r.gl_seq, arena.floor_seq) > arena.ready_seq):
                later.append(r)
            else:
                reqs.append(r)
        self._admitted_grid = later
        if not reqs:
            return
        items = np.concatenate([r.items for r in reqs]) if len(reqs) > 1 else reqs[0].items.copy()
        gl_n = np.concatenate([r.gl_n for r in reqs]) if len(reqs) > 1 else reqs[0].gl_n
        count = len(items)
        is_gl = gl_n > 0
        gl_len = 0
        gl_x = gl_w = gl_p = None
        if np.any(is_gl):
            items["gl_offset"][is_gl] = arena.offsets(gl_n[is_gl])
            gl_x, gl_w, gl_p, gl_len = arena.x, arena.w, arena.p, arena.used
        with torch.cuda.stream(slot.stream):
            # the host planner reads the records where they are; its outputs
            # (descriptors, tiles, per-integral tile ranges) go into one pinned
            # staging buffer and reach the device in one copy
            base = ctypes.c_void_p(items.ctypes.data)
            desc_bytes = ctypes.c_int64()
            n_tiles = ctypes.c_int64()
            call("tq_batch_plan_sizes", base, count, self.dt, self.tile_points, ctypes.byref(desc_bytes),
                 ctypes.byref(n_tiles))
            nt = max(int(n_tiles.value), 1)
            o_tiles = _aligned(desc_bytes.value)
            o_begin = o_tiles + _aligned(nt * TILE_BYTES)
            o_count = o_begin + _aligned(count * 8)
            total = o_count + count * 8
            sh, sd = slot.stage.reserve(total)
            hp, dp = sh.data_ptr(), sd.data_ptr()
            call("tq_batch_plan", base, count, self.dt, self.tile_points, ctypes.c_void_p(hp),
                 ctypes.c_void_p(hp + o_tiles), ctypes.c_void_p(hp + o_begin), ctypes.c_void_p(hp + o_count),
                 slot.class_offsets.ctypes.data_as(ctypes.c_void_p))
            sd[:total].copy_(sh[:total], non_blocking=True)
            slot.ensure_outputs(count, nt)
            call("tq_batch_run", self.dt, ctypes.c_void_p(dp), count, ctypes.c_void_p(dp + o_tiles),
                 int(n_tiles.value), ctypes.c_void_p(dp + o_begin), ctypes.c_void_p(dp + o_count), ptr(gl_x),
                 ptr(gl_w), ptr(gl_p), gl_len, ptr(slot.part), ptr(slot.out), ptr(slot.out64),
                 slot.class_offsets.ctypes.data_as(ctypes.c_void_p), slot.stream.cuda_stream)
            slot.host_out[:count].copy_(slot.out64[:count], non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(slot.stream)
        slot.event = ev
        slot.requests = reqs
        slot.gl = (gl_x, gl_w, gl_p)
        self.batches_launched += 1

    def _advance_vegas(self):
        from .integration.vegas import VEGAS

        finished = 0
        for r in list(self._vegas):
            t = r.ticket
            if t.cancelled:
                if r.event is None or r.event.query():
                    self._vegas.remove(r)
                    self._release(t)
                    t._finish()
                    finished += 1
                continue
            if r.wait and not r.event.query():
                continue
            try:
                with torch.cuda.stream(self._vegas_stream(r)):
                    if r.integrator is None:
                        a = r.args
                        dom = a["domain"] if a["domain"] is not None else [[-1.0, 1.0]] * a["dim"]
                        dom = torch.as_tensor(dom, dtype=self.dtype, device=self.device)
                        v = VEGAS()
                        v._setup(a["fn"], a["dim"], a["N"], dom, a["seed"], None, a["use_grid_improve"],
                                 a["eps_rel"], a["eps_abs"], a["max_iterations"], None, a["N_intervals"],
                                 a["N_strat"], a["warm_start"])
                        r.integrator = v
                        r.steps = v._steps(a["use_warmup"])
                    try:
                        r.wait = next(r.steps)
                        if r.wait:
                            if r.event is None:
                                r.event = torch.cuda.Event()
                            r.event.record(r.stream)
                        continue
                    except StopIteration:
                        value = float(r.integrator._finish())
                self._vegas.remove(r)
                self._release(t)
                t.state = r.integrator.state_dict()
                t._finish(value=value)
                self.completed.append(t)
                finished += 1
            except Exception as e:  # the request fails, the server goes on
                self._vegas.remove(r)
                self._release(t)
                t._finish(error=e)
                finished += 1
        return finished

    def _vegas_stream(self, r):
        # every VEGAS request shares one stream: with the batch slots' and the
        # table stream that keeps the server within the device's 4 hardware
        # queues (more streams would share queues and add false dependencies)
        if self._vegas_shared_stream is None:
            self._vegas_shared_stream = torch.cuda.Stream(device=self.device)
        r.stream = self._vegas_shared_stream
        return r.stream

    # -- driving ------------------------------------------------------------
    def run_until_idle(self, timeout=None):
        """Step until nothing is waiting or in flight."""
        t0 = time.perf_counter()
        while self.pending():
            if not self.step():
                time.sleep(0)
            if timeout is not None and time.perf_counter() - t0 > timeout:
                raise TimeoutError("server did not drain in time")

    def start(self):
        """Run the scheduling loop in a background thread."""
        if self._thread is not None:
            return
        self._stop.clear()

        def loop():
            torch.cuda.set_device(self.device)
            while not self._stop.is_set():
                if not self.pending():
                    time.sleep(50e-6)
                    continue
                self.step()

        self._thread = threading.Thread(target=loop, name="tq-serve", daemon=True)
        self._thread.start()

    def stop(self):
        if self._thread is None:
            return
        self._stop.set()
        self._thread.join()
        self._thread = None

    def __enter__(self):
        self.start()
        return self

    def __exit__(self, *exc):
        self.stop()
        return False


def latency_stats(tickets):
    """p50 / p99 / mean latency in ms over the finished integrals (a ticket of
    submit_many counts once per integral, all with its latency)."""
    lat = sorted(x for t in tickets if t.latency is not None and t.error is None
                 for x in [t.latency * 1e3] * t.count)
    if not lat:
        return {"count": 0}
    return {
        "count": len(lat),
        "p50_ms": lat[len(lat) // 2],
        "p99_ms": lat[min(len(lat) - 1, int(0.99 * len(lat)))],
        "mean_ms": sum(lat) / len(lat),
        "max_ms": lat[-1],
    }


def poisson_arrivals(rate, count, seed=0):
    """Arrival offsets (s) of a Poisson process of `rate` requests/s."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, count))


def serve_open_loop(server, requests, rate, seed=0):
    """Replay `requests` (callables taking (server, arrival) and returning a
    Ticket) at Poisson arrival times in this thread, stepping the server
    between arrivals; returns the tickets."""
    offs = poisson_arrivals(rate, len(requests), seed)
    t0 = time.perf_counter()
    tickets = []
    i = 0
    while i < len(requests) or server.pending():
        now = time.perf_counter() - t0
        while i < len(requests) and offs[i] <= now:
            tickets.append(requests[i](server, t0 + offs[i]))
            i += 1
        server.step()
    return tickets
