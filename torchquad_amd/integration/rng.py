"""Random number generator (reference: torchquad/integration/rng.py:5-152, torch path :77-125).

Uniform numbers come from the library's counter-based Philox-4x32-10
stream on the GPU: the object holds a 64-bit key (the seed) and a counter
position that advances by one per generated row, so draws are reproducible,
independent of launch shape and never touch torch's global generator state
for the numbers themselves.  As in the reference, a given seed is also
applied to torch's global generator (rng.py:121-124), and seed=None draws a
key from torch's global CPU generator (so torch.manual_seed controls it).
"""
import numpy as np
import torch

from .. import _ops


class RNG:
    """Seedable uniform [0, 1) generator for the torch backend."""

    def __init__(self, backend="torch", seed=None, torch_save_state=False):
        if backend not in ("torch", None):
            raise ValueError(f"Unsupported numerical backend: {backend}. torchquad_amd uses the torch backend.")
        if seed is None:
            seed = int(torch.randint(0, 2**62, (1,), generator=None).item())
        elif not torch_save_state:
            torch.random.manual_seed(seed)
        self._seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        self._offset = 0
        self.uniform = self._uniform

    @property
    def key(self):
        return self._seed

    def reserve(self, rows):
        """Advance the stream by `rows` samples; returns the first counter."""
        first = self._offset
        self._offset += int(rows)
        return first

    def _uniform(self, size, dtype=None, device=None):
        """Uniform numbers in [0, 1) of the given shape (rng.py:127-139)."""
        if dtype is None:
            dtype = torch.get_default_dtype()
        if device is None:
            device = torch.device("cuda", torch.cuda.current_device()) if torch.cuda.is_available() else torch.device("cpu")
        size = [int(s) for s in (size if isinstance(size, (list, tuple)) else [size])]
        if len(size) == 0:
            rows, cols = 1, 1
        elif len(size) == 1:
            rows, cols = size[0], 1
        else:
            rows, cols = int(np.prod(size[:-1])), size[-1]
        first = self.reserve(rows)
        out = _ops.uniform(self._seed, first, rows, cols, dtype, torch.device(device))
        return out.reshape(size)

    def jax_get_key(self):  # pragma: no cover - JAX is not a backend of this package
        raise NotImplementedError("JAX is not supported by torchquad_amd")

    def jax_set_key(self, key):  # pragma: no cover
        raise NotImplementedError("JAX is not supported by torchquad_amd")
