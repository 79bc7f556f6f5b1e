"""VEGAS stratification (reference: torchquad/integration/vegas_stratification.py:6-165).

Damped counts dh, per-cube sums JF/JF2 and the per-cube sample counts live
on the GPU in FP64; attributes return the requested dtype.
"""
import torch

from .. import _vegas_dev
from .._lib import call, ptr
from .rng import RNG


def _cuda():
    return torch.device("cuda", torch.cuda.current_device()) if torch.cuda.is_available() else torch.device("cpu")


def rng_is_native(rng):
    """True when rng draws from the library's Philox stream (not a subclass
    that post-processes the numbers, e.g. vegas_test.py:143-156)."""
    return isinstance(rng, RNG) and getattr(rng.uniform, "__func__", None) is RNG._uniform


class VEGASStratification:
    """Hypercube stratification of VEGAS+ (arXiv:2009.05112, section III)."""

    def __init__(self, N_increment, dim, rng, backend, dtype, beta=0.75, _device=None):
        self.rng = rng
        self.dim = dim
        self.N_strat = _device.n_strat if _device is not None else _vegas_dev.n_strat_for(N_increment, dim)
        self.beta = beta
        self.N_cubes = self.N_strat**self.dim
        self.V_cubes = (1.0 / self.N_strat) ** self.dim
        self.dtype = dtype
        self.backend = backend
        if _device is None:
            _device = _vegas_dev.VegasDevice(dim, 2, self.N_strat, dtype, _cuda(), beta=beta)
        self._dev = _device
        self._last_nevals = None

    @property
    def dh(self):
        return self._dev.dh.to(self.dtype)

    @property
    def JF(self):
        return self._dev.jf_cube.to(self.dtype)

    @property
    def JF2(self):
        return self._dev.jf2_cube.to(self.dtype)

    @property
    def strat_counts(self):
        return self._dev.nh.to(self.dtype)

    def _set_counts(self, nevals):
        """offsets / cube index for given per-cube counts."""
        dev = self._dev
        nevals = nevals.to(device=dev.device, dtype=torch.int64).contiguous()
        call("tq_vegas_nh", dev.struct, 0, -1, 0, ptr(nevals), dev.stream())
        total = int(dev.offsets[-1].item())
        if total > dev.max_samples:
            dev.ensure_capacity(total)
            call("tq_vegas_nh", dev.struct, 0, -1, 0, ptr(nevals), dev.stream())
        return total

    def accumulate_weight(self, nevals, weight_all_cubes):
        """Per-cube sums JF, JF2 of the sorted sample weights (vegas_stratification.py:46-70)."""
        total = self._set_counts(nevals)
        w = weight_all_cubes.detach().to(device=self._dev.device, dtype=torch.float64).contiguous()
        if w.shape[0] != total:
            raise ValueError(f"{w.shape[0]} weights for {total} samples")
        call("tq_vegas_strat_accumulate", self._dev.struct, ptr(w), 0, self._dev.stream())
        return self.JF, self.JF2

    def update_DH(self):
        """dh proportional to the per-cube variance^beta, normalised (vegas_stratification.py:72-90)."""
        call("tq_vegas_update_dh", self._dev.struct, 1, self._dev.stream())
        call("tq_vegas_normalize_dh", self._dev.struct, 1, self._dev.stream())

    def get_NH(self, nevals_exp):
        """nh = max(2, floor(dh * nevals_exp)) per cube, EQ 44 (vegas_stratification.py:92-103)."""
        dev = self._dev
        dev.ensure_capacity(dev.capacity_for(nevals_exp))
        call("tq_vegas_nh", dev.struct, int(nevals_exp), -1, 0, None, dev.stream())
        return dev.nh.clone()

    def _get_indices(self, idx):
        """Cube index -> per-axis digits, dim 0 fastest (vegas_stratification.py:105-138)."""
        points = torch.reshape(idx, [idx.shape[0], 1])
        strides = self.N_strat ** torch.arange(self.dim, device=idx.device)
        points = torch.div(points, strides, rounding_mode="floor")
        points[:, :-1] = points[:, :-1] - self.N_strat * points[:, 1:]
        return points

    def get_Y(self, nevals):
        """Stratified samples in [0,1)^dim (vegas_stratification.py:140-165)."""
        total = self._set_counts(nevals)
        dev = self._dev
        y = torch.empty((total, self.dim), dtype=dev.sample_dtype, device=dev.device)
        if total == 0:
            return y
        if rng_is_native(self.rng):
            first = self.rng.reserve(total)
            call("tq_vegas_strat_y", dev.struct, self.rng.key, first, None, ptr(y), dev.stream())
        else:
            u = self.rng.uniform(size=[total, self.dim], dtype=dev.sample_dtype)
            u = u.to(device=dev.device, dtype=dev.sample_dtype).contiguous()
            call("tq_vegas_strat_y", dev.struct, 0, 0, ptr(u), ptr(y), dev.stream())
        return y
