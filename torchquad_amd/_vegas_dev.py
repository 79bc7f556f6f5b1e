"""Device-resident VEGAS state (tq_vegas_state of include/torchquad_amd.h).

All buffers are torch tensors owned by a VegasDevice object; the struct
handed to the C ABI holds their device pointers.  Map and stratification
state are FP64; samples and the integrand use the sample dtype.
"""
import ctypes
import math

import torch

from . import _lib
from ._lib import call, ptr, stream_ptr

SB = 1024  # TQ_VEGAS_BLOCK_SAMPLES
MAX_ITERS = 4096  # TQ_VEGAS_MAX_ITERS
HIST_LDS = 768


VegasState = _lib.VegasState


def n_strat_for(N_increment, dim):
    """stratification steps per axis, EQ 41 (vegas_stratification.py:27-28)."""
    n = int((N_increment / 4.0) ** (1.0 / dim))
    return 1000 if n > 1000 else n


class VegasDevice:
    """Owns every buffer of one VEGAS map + stratification."""

    def __init__(self, dim, n_intervals, n_strat, sample_dtype, device, alpha=0.5, beta=0.75, max_samples=SB):
        if not torch.cuda.is_available() or torch.device(device).type != "cuda":
            raise _lib.NativeUnavailableError("VEGAS runs on a ROCm GPU only")
        self.dim = int(dim)
        self.n_intervals = int(n_intervals)
        self.n_strat = int(max(n_strat, 1))
        self.n_cubes = self.n_strat**self.dim
        self.v_cubes = (1.0 / self.n_strat) ** self.dim
        self.sample_dtype = sample_dtype
        self.device = torch.device(device)
        self.alpha = float(alpha)
        self.beta = float(beta)
        f64 = dict(dtype=torch.float64, device=self.device)
        i64 = dict(dtype=torch.int64, device=self.device)
        d, ni, nc = self.dim, self.n_intervals, self.n_cubes
        self.x_edges = torch.zeros((d, ni + 1), **f64)
        self.dx_edges = torch.zeros((d, ni), **f64)
        self.weights = torch.zeros((d, ni), **f64)
        self.counts = torch.zeros((d, ni), **f64)
        self.dh = torch.zeros(nc, **f64)
        self.jf_cube = torch.zeros(nc, **f64)
        self.jf2_cube = torch.zeros(nc, **f64)
        self.nh = torch.zeros(nc, **i64)
        self.offsets = torch.zeros(nc + 1, **i64)
        nb_scan = (nc + 2047) // 2048
        self.scan_sums = torch.zeros(nb_scan + 1, **i64)
        self.iter_stats = torch.zeros((MAX_ITERS, 4), **f64)
        self.flags = torch.zeros(4, dtype=torch.int32, device=self.device)
        self.max_samples = 0
        self.max_blocks = 0
        self.ensure_capacity(max(int(max_samples), SB))
        self.init()

    def ensure_capacity(self, n_samples):
        """Grow the per-sample and per-block buffers to hold n_samples samples."""
        n_samples = int(n_samples)
        if n_samples <= self.max_samples and self.max_blocks > 0:
            return
        cap = max(n_samples, int(self.max_samples * 1.25), SB)
        blocks = max((cap + SB - 1) // SB, 4)
        d, ni = self.dim, self.n_intervals
        W = d * ni
        f64 = dict(dtype=torch.float64, device=self.device)
        self.jf = torch.zeros(cap, **f64)
        self.hist_w = torch.zeros(max(blocks * W, 5 * d * (ni + 1)), **f64)
        self.hist_c = torch.zeros(blocks * W, dtype=torch.int32, device=self.device)
        self.chunk_cube = torch.zeros(blocks + 1, dtype=torch.int64, device=self.device)
        nb_strat = (self.n_cubes + 255) // 256
        self.block_sums = torch.zeros((max(blocks, nb_strat), 4), **f64)
        self.max_samples = cap
        self.max_blocks = blocks
        self._struct = None

    @property
    def struct(self):
        if self._struct is None:
            s = VegasState()
            s.dim = self.dim
            s.n_intervals = self.n_intervals
            s.n_strat = self.n_strat
            s.sample_dtype = _lib.dtype_code(self.sample_dtype)
            s.n_cubes = self.n_cubes
            s.max_samples = self.max_samples
            s.max_blocks = self.max_blocks
            s.v_cubes = self.v_cubes
            s.alpha = self.alpha
            s.beta = self.beta
            for name in ("x_edges", "dx_edges", "weights", "counts", "dh", "jf_cube", "jf2_cube", "nh",
                         "offsets", "chunk_cube", "jf", "hist_w", "hist_c", "block_sums", "scan_sums",
                         "iter_stats", "flags"):
                setattr(s, name, getattr(self, name).data_ptr())
            self._struct = s
        return ctypes.byref(self._struct)

    _BUFFERS = ("x_edges", "dx_edges", "weights", "counts", "dh", "jf_cube", "jf2_cube", "nh", "offsets",
                "scan_sums", "iter_stats", "flags", "jf", "hist_w", "hist_c", "chunk_cube", "block_sums")

    def reset(self):
        """Bring a used state back to the freshly built one (every buffer zero, then the map and
        stratification initialised) with one multi-tensor fill instead of new allocations."""
        torch._foreach_zero_([getattr(self, name) for name in self._BUFFERS])
        self.init()

    def stream(self):
        return stream_ptr(self.device)

    def init(self):
        call("tq_vegas_init", self.struct, self.stream())

    def capacity_for(self, n_exp):
        """Upper bound of sum(nh) for an expected count (floor + clip at 2)."""
        return int(n_exp) + 3 * self.n_cubes + SB

    # -- checks at synchronisation points -------------------------------------
    def read_flags(self):
        return [int(v) for v in self.flags.tolist()]

    def clear_flags(self):
        self.flags.zero_()


def smooth(weights, counts, alpha):
    """_smooth_map on device tensors; returns (smoothed, not_updatable)."""
    weights = weights.to(torch.float64).contiguous()
    counts = counts.to(torch.float64).contiguous()
    d, n = weights.shape
    out = torch.zeros_like(weights)
    scratch = torch.zeros((d, 2 * n), dtype=torch.float64, device=weights.device)
    flag = torch.zeros(1, dtype=torch.int32, device=weights.device)
    call("tq_vegas_smooth", int(d), int(n), ptr(weights), ptr(counts), float(alpha), ptr(out), ptr(scratch),
         ptr(flag), stream_ptr(weights.device))
    return out, bool(flag.item())
