"""VEGAS adaptive map (reference: torchquad/integration/vegas_map.py:9-261).

Map state (edges, widths, accumulated weights and counts) lives on the GPU
in FP64.  The public attributes return tensors in the requested dtype, as
the reference's do; the fused VEGAS pipeline works on the FP64 buffers.
"""
import warnings

import torch

from .. import _vegas_dev
from .._lib import call, ptr
from ..utils.set_log_level import logger


def _cuda():
    return torch.device("cuda", torch.cuda.current_device()) if torch.cuda.is_available() else torch.device("cpu")


class VEGASMap:
    """Adaptive 1-D piecewise-linear map per axis (VEGAS+, arXiv:2009.05112)."""

    def __init__(self, N_intervals, dim, backend, dtype, alpha=0.5, _device=None) -> None:
        self.dim = dim
        self.N_intervals = N_intervals
        self.alpha = alpha
        self.backend = backend
        self.dtype = dtype
        if _device is None:
            _device = _vegas_dev.VegasDevice(dim, N_intervals, 1, dtype, _cuda(), alpha=alpha)
        self._dev = _device

    # -- state, in the requested dtype ----------------------------------------
    @property
    def x_edges(self):
        return self._dev.x_edges.to(self.dtype)

    @property
    def dx_edges(self):
        return self._dev.dx_edges.to(self.dtype)

    @property
    def weights(self):
        return self._dev.weights.to(self.dtype)

    @property
    def counts(self):
        return self._dev.counts.to(torch.int64)

    def _y(self, y):
        return y.to(self._dev.sample_dtype).contiguous()

    def _eval(self, y, ids=False, off=False, x=False, jac=False):
        y = self._y(y)
        n = y.shape[0]
        dev = self._dev
        o_ids = torch.empty((n, self.dim), dtype=torch.int64, device=y.device) if ids else None
        o_off = torch.empty_like(y) if off else None
        o_x = torch.empty_like(y) if x else None
        o_jac = torch.empty((n,), dtype=y.dtype, device=y.device) if jac else None
        call("tq_vegas_map_eval", dev.struct, ptr(y), n, ptr(o_ids), ptr(o_off), ptr(o_x), ptr(o_jac), dev.stream())
        return o_ids, o_off, o_x, o_jac

    def get_X(self, y):
        """Mapped points x = e[ID] + dx[ID] * offset, EQ 9 (vegas_map.py:44-58)."""
        return self._eval(y, x=True)[2]

    def get_Jac(self, y):
        """Jacobian prod_k N_intervals * dx[ID_k], EQ 12 (vegas_map.py:60-74)."""
        return self._eval(y, jac=True)[3]

    def _get_interval_ID(self, y):
        """floor(y * N_intervals) as int64, EQ 10 (vegas_map.py:76-85)."""
        return self._eval(y, ids=True)[0]

    def _get_interval_offset(self, y):
        """y * N_intervals - floor(.), EQ 11 (vegas_map.py:87-97)."""
        return self._eval(y, off=True)[1]

    def accumulate_weight(self, y, jf_vec2):
        """Add (f J)^2 and counts per bin (vegas_map.py:99-111)."""
        y = self._y(y)
        jf2 = jf_vec2.detach().to(self._dev.sample_dtype).contiguous()
        call("tq_vegas_map_accumulate", self._dev.struct, ptr(y), ptr(jf2), y.shape[0], self._dev.stream())

    @staticmethod
    def _smooth_map(weights, counts, alpha):
        """Smoothed, range-compressed weights, EQ 18-22 (vegas_map.py:113-172);
        None when some axis has only zero weights."""
        dtype = weights.dtype
        out, not_updatable = _vegas_dev.smooth(weights, counts, alpha)
        if not_updatable:
            return None
        return out.to(dtype)

    def _reset_weight(self):
        """vegas_map.py:174-183"""
        self._dev.weights.zero_()
        self._dev.counts.zero_()

    def update_map(self):
        """Redistribute the edges from the accumulated weights (vegas_map.py:185-261)."""
        dev = self._dev
        before = dev.read_flags()
        call("tq_vegas_map_update_accumulated", dev.struct, dev.stream())
        after = dev.read_flags()
        report_map_flags(before, after)


def report_map_flags(before, after):
    """Warn / raise for map update events recorded on the device."""
    if after[0] > before[0]:
        msg = "Cannot update the VEGASMap. This can happen with an integrand which evaluates to zero everywhere."
        logger.warning(msg)
        warnings.warn(msg, RuntimeWarning)
    if after[1] > before[1]:
        msg = "Some calculated VEGASMap edges were infinite"
        logger.warning(msg)
        warnings.warn(msg, RuntimeWarning)
    if after[2] > before[2]:
        raise RuntimeError("Could not replace all infinite edges")
