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
   (``tq_batch_plan``), asynchronous upload from pinned buffers, one
   persistent kernel plus the per-integral reduction (``tq_batch_run``),
   asynchronous download of the results -- on one of ``max_inflight``
   streams, so a batch is planned and launched while earlier ones still
   run;
4. advances each resident VEGAS request by one iteration on its own
   stream.  Map and stratification state stay on the device between
   iterations; the every-fifth-iteration abort check runs only once the
   request's event shows the device has finished those iterations.

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
import threading
import time
from collections import OrderedDict, deque

import numpy as np
import torch

from . import _lib
from ._lib import NativeUnavailableError, call, lib, ptr
from .batch import RULE_NAMES, n_per_axis
from .gl_nodes import fp32_pairs, gauss_legendre
from .integrands import BuiltinIntegrand

ITEM_BYTES = ctypes.sizeof(_lib.BatchItem)
TILE_BYTES = ctypes.sizeof(_lib.BatchTile)


class AdmissionError(RuntimeError):
    """The request can never fit the server's HBM budget."""


class CancelledError(RuntimeError):
    """The request was cancelled before its result was delivered."""


class Ticket:
    """Handle of one request: wait for its result, or cancel it."""

    _ids = itertools.count()

    def __init__(self, kind, nbytes, evals):
        self.id = next(Ticket._ids)
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
        self._event = threading.Event()

    def done(self):
        return self._event.is_set()

    @property
    def latency(self):
        """Seconds from arrival to delivery (None until done)."""
        return None if self.t_done is None else self.t_done - self.t_arrival

    def result(self, timeout=None):
        """The integral (float64 Python float); raises the request's error."""
        if not self._event.wait(timeout):
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
        self._event.set()


class _GridRequest:
    __slots__ = ("ticket", "item", "gl_n", "gl_entry")

    def __init__(self, ticket, item, gl_n):
        self.ticket = ticket
        self.item = item
        self.gl_n = gl_n
        self.gl_entry = None  # (nodes, weights, event, FP32 pairs) once its table is started


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
        self.items = _Pinned(device)
        self.desc = _Pinned(device)
        self.tiles = _Pinned(device)
        self.tb = _Pinned(device)
        self.tc = _Pinned(device)
        self.part = torch.empty(0, dtype=torch.float64, device=device)
        self.counter = torch.zeros(_lib.BATCH_COUNTER_WORDS, dtype=torch.int32, device=device)
        self.class_offsets = np.zeros(4, dtype=np.int64)
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
    """

    def __init__(self, device=None, dtype=torch.float32, hbm_budget=None, tile_points=8192, max_inflight=3,
                 max_vegas=8, max_batch=8192):
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
        self._gl_cache = OrderedDict()
        self._gl_cache_bytes = 0
        self._gl_streams = []  # tables of new n are generated round-robin on these
        self._gl_next = 0
        self.gl_cache_limit = 256 << 20
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
        item = _lib.BatchItem()
        item.rule, item.family, item.dim, item.n = code, d.family, dim, int(n)
        item.gl_offset = 0
        for k in range(dim):
            item.domain[2 * k] = dom[k, 0]
            item.domain[2 * k + 1] = dom[k, 1]
        item.c0, item.c1, item.c2, item.c3 = d.c0, d.c1, d.c2, d.c3
        for k in range(_lib.TQ_MAX_DIM):
            item.a[k] = d.a[k]
            item.u[k] = d.u[k]
        desc_bytes = ctypes.c_int64()
        n_tiles = ctypes.c_int64()
        call("tq_batch_plan_sizes", ctypes.byref(item), 1, self.dt, self.tile_points, ctypes.byref(desc_bytes),
             ctypes.byref(n_tiles))
        gl_n = int(n) if code == _lib.RULE_GAUSS_LEGENDRE else 0
        gl_bytes = (24 if self.dt == _lib.TQ_F32 else 16) * gl_n  # FP64 table (+ FP32 pairs) in the batch
        nbytes = (desc_bytes.value + n_tiles.value * (TILE_BYTES + 8) + 16 + 16 + ITEM_BYTES + gl_bytes)
        t = Ticket("grid", nbytes, int(n) ** dim)
        t.rule = code
        if arrival is not None:
            t.t_arrival = arrival
        self._enqueue(_GridRequest(t, item, gl_n))
        return t

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
                vals = slot.host_out[: len(slot.requests)].tolist()
                for r, v in zip(slot.requests, vals):
                    self._release(r.ticket)
                    if not r.ticket.cancelled:
                        r.ticket._finish(value=v)
                    else:
                        r.ticket._finish()
                    self.completed.append(r.ticket)
                finished += len(slot.requests)
                slot.requests = None
                slot.gl = None
                slot.event = None
        return finished

    def _admit(self):
        with self._lock:
            while self._waiting:
                r = self._waiting[0]
                if self.reserved + r.ticket.nbytes > self.hbm_budget:
                    break
                if isinstance(r, _VegasRequest) and len(self._vegas) >= self.max_vegas:
                    break
                if isinstance(r, _GridRequest) and len(self._admitted_grid) >= self.max_batch:
                    break
                self._waiting.popleft()
                self.reserved += r.ticket.nbytes
                self.peak_reserved = max(self.peak_reserved, self.reserved)
                r.ticket.t_admit = time.perf_counter()
                if isinstance(r, _GridRequest):
                    self._admitted_grid.append(r)
                else:
                    self._vegas.append(r)

    def _gl_table(self, n):
        """The n-point Gauss-Legendre table entry (nodes, weights, ready event,
        FP32 (node, weight) pairs or None for a float64 server).

        A new n starts its table on a side stream, so a request whose table
        is still being generated waits for a later batch instead of delaying
        the requests packed with it.  Requests keep a reference to their
        entry, so evicting it from the cache (least recently used first,
        beyond gl_cache_limit bytes) never takes a table from a request that
        still needs it.
        """
        e = self._gl_cache.get(n)
        if e is not None:
            self._gl_cache.move_to_end(n)
            return e
        if not self._gl_streams:
            self._gl_streams = [torch.cuda.Stream(device=self.device) for _ in range(4)]
        stream = self._gl_streams[self._gl_next % len(self._gl_streams)]
        self._gl_next += 1
        with torch.cuda.stream(stream):
            x, w = gauss_legendre(n, device=self.device)
            p32 = fp32_pairs(x, w) if self.dt == _lib.TQ_F32 else None
            ev = torch.cuda.Event()
            ev.record(stream)
        e = (x, w, ev, p32)
        per = 24 if p32 is not None else 16
        self._gl_cache[n] = e
        self._gl_cache_bytes += per * n
        while self._gl_cache_bytes > self.gl_cache_limit and len(self._gl_cache) > 1:
            old_n, _ = self._gl_cache.popitem(last=False)
            self._gl_cache_bytes -= per * old_n
        return e

    def _launch_grid(self):
        if not self._admitted_grid:
            return
        for r in self._admitted_grid:  # start tables even while every slot is busy
            if r.gl_n and r.gl_entry is None:
                r.gl_entry = self._gl_table(r.gl_n)
        slot = None
        for k in range(len(self._slots)):
            s = self._slots[(self._next_slot + k) % len(self._slots)]
            if not s.busy():
                slot = s
                self._next_slot = (self._next_slot + k + 1) % len(self._slots)
                break
        if slot is None:
            return
        reqs, later, ready = [], [], {}
        for r in self._admitted_grid:
            if r.gl_n:
                if not r.gl_entry[2].query():
                    later.append(r)
                    continue
                ready.setdefault(r.gl_n, r.gl_entry)
            reqs.append(r)
        self._admitted_grid = later
        if not reqs:
            return
        count = len(reqs)
        # Gauss-Legendre tables of the batch (ready), concatenated on the slot's stream
        gl_x = gl_w = gl_p = None
        ns = sorted({r.gl_n for r in reqs if r.gl_n})
        offsets, tabs, off = {}, [], 0
        for n in ns:
            offsets[n] = off
            tabs.append(ready[n])
            off += n
        with torch.cuda.stream(slot.stream):
            if ns:
                for x, w, ev, p32 in tabs:
                    slot.stream.wait_event(ev)
                    x.record_stream(slot.stream)
                    w.record_stream(slot.stream)
                    if p32 is not None:
                        p32.record_stream(slot.stream)
                gl_x = torch.cat([t[0] for t in tabs])
                gl_w = torch.cat([t[1] for t in tabs])
                if self.dt == _lib.TQ_F32:
                    gl_p = torch.cat([t[3] for t in tabs])
                for r in reqs:
                    if r.gl_n:
                        r.item.gl_offset = offsets[r.gl_n]
            ih, idev = slot.items.reserve(count * ITEM_BYTES)
            base = ih.data_ptr()
            for i, r in enumerate(reqs):
                ctypes.memmove(base + i * ITEM_BYTES, ctypes.addressof(r.item), ITEM_BYTES)
            desc_bytes = ctypes.c_int64()
            n_tiles = ctypes.c_int64()
            call("tq_batch_plan_sizes", ctypes.c_void_p(base), count, self.dt, self.tile_points,
                 ctypes.byref(desc_bytes), ctypes.byref(n_tiles))
            nt = max(int(n_tiles.value), 1)
            dh, dd = slot.desc.reserve(desc_bytes.value)
            th, td = slot.tiles.reserve(nt * TILE_BYTES)
            bh, bd = slot.tb.reserve(count * 8)
            ch, cd = slot.tc.reserve(count * 8)
            call("tq_batch_plan", ctypes.c_void_p(base), count, self.dt, self.tile_points, ptr(dh), ptr(th), ptr(bh),
                 ptr(ch), slot.class_offsets.ctypes.data_as(ctypes.c_void_p))
            dd[: desc_bytes.value].copy_(dh[: desc_bytes.value], non_blocking=True)
            td[: nt * TILE_BYTES].copy_(th[: nt * TILE_BYTES], non_blocking=True)
            bd[: count * 8].copy_(bh[: count * 8], non_blocking=True)
            cd[: count * 8].copy_(ch[: count * 8], non_blocking=True)
            slot.ensure_outputs(count, nt)
            call("tq_batch_run", self.dt, ptr(dd), count, ptr(td), int(n_tiles.value), ptr(bd), ptr(cd), ptr(gl_x),
                 ptr(gl_w), ptr(gl_p), ptr(slot.part), ptr(slot.counter), slot.counter.numel(), ptr(slot.out), ptr(slot.out64), 0,
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
        if r.stream is None:
            r.stream = torch.cuda.Stream(device=self.device)
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
    """p50 / p99 / mean latency in ms over finished tickets."""
    lat = sorted(t.latency * 1e3 for t in tickets if t.latency is not None and t.error is None)
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
