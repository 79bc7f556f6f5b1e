"""Built-in integrand families that the fused HIP kernels evaluate in-register.

Passing one of these objects as ``fn`` to an integrator selects the fused
path: points are never written to memory.  Each object is also an ordinary
callable on a ``[N, dim]`` tensor (its mathematical definition in torch), so
it can be handed to any code that expects a torchquad integrand, and the
generic user-integrand path can be compared against the fused one.

Families (parameters are per-axis vectors of length ``dim``):

* ``GaussianPeak(a, u)``      f = exp(-sum a_k^2 (x_k - u_k)^2)       (Genz "Gaussian")
* ``Oscillatory(c, w)``       f = cos(2 pi w + sum c_k x_k)           (Genz "oscillatory")
* ``ProductPeak(a, u)``       f = prod 1 / (a_k^-2 + (x_k - u_k)^2)   (Genz "product peak")
* ``ExpLinear(a, c)``         f = exp(c + sum a_k x_k)
* ``SinSum(a, u)``            f = sum sin(a_k x_k + u_k)
* ``DoubleGaussian(width, centers)``
  f = 1/2 sum_{c in centers} (width sqrt(pi))^-dim exp(-sum (x_k - c)^2 / width^2)
* ``Constant(value)``         f = value
"""
import math

import numpy as np
import torch

from . import _lib


def _vec(v, dim, default):
    if v is None:
        return np.full(dim, float(default), dtype=np.float64)
    if isinstance(v, torch.Tensor):
        v = v.detach().to("cpu", torch.float64).numpy()
    arr = np.asarray(v, dtype=np.float64).reshape(-1)
    if arr.size == 1 and dim > 1:
        arr = np.full(dim, float(arr[0]))
    if arr.size != dim:
        raise ValueError(f"parameter has {arr.size} entries, expected dim={dim}")
    return arr


class BuiltinIntegrand:
    """Base class: a family id, a dimension and float64 parameters."""

    family = None
    name = "builtin"

    def __init__(self, dim, scale=1.0):
        self.dim = int(dim)
        if self.dim < 1:
            raise ValueError("dim must be >= 1")
        self.scale = float(scale)

    # -- fused-kernel descriptor ------------------------------------------
    def _fill(self, d):  # pragma: no cover - overridden
        raise NotImplementedError

    def descriptor(self, dim=None):
        dim = self.dim if dim is None else int(dim)
        if dim != self.dim:
            raise ValueError(f"{self.name} was built for dim={self.dim}, called with dim={dim}")
        if dim > _lib.TQ_MAX_DIM:
            raise ValueError(f"fused integrands support dim <= {_lib.TQ_MAX_DIM}")
        d = _lib.Integrand()
        d.family = self.family
        d.dim = dim
        d.c0 = self.scale
        self._fill(d)
        return d

    # -- torch definition -------------------------------------------------
    def torch_eval(self, x):  # pragma: no cover - overridden
        raise NotImplementedError

    def __call__(self, x):
        if x.shape[-1] != self.dim:
            raise ValueError(f"{self.name} expects points of dim {self.dim}, got {x.shape[-1]}")
        return self.scale * self.torch_eval(x)

    def _t(self, arr, x):
        return torch.as_tensor(arr, dtype=x.dtype, device=x.device)

    def __repr__(self):
        return f"{type(self).__name__}(dim={self.dim})"


class GaussianPeak(BuiltinIntegrand):
    family = _lib.FAM_GAUSSIAN_PEAK
    name = "GaussianPeak"

    def __init__(self, a, u, dim=None, scale=1.0):
        a_arr = np.asarray(a.detach().cpu() if isinstance(a, torch.Tensor) else a, dtype=np.float64).reshape(-1)
        dim = a_arr.size if dim is None else dim
        super().__init__(dim, scale)
        self.a = _vec(a, self.dim, 1.0)
        self.u = _vec(u, self.dim, 0.0)

    def _fill(self, d):
        for k in range(self.dim):
            d.a[k] = self.a[k]
            d.u[k] = self.u[k]

    def torch_eval(self, x):
        a, u = self._t(self.a, x), self._t(self.u, x)
        return torch.exp(-torch.sum((a * (x - u)) ** 2, dim=-1))


class Oscillatory(BuiltinIntegrand):
    family = _lib.FAM_OSCILLATORY
    name = "Oscillatory"

    def __init__(self, c, w=0.0, dim=None, scale=1.0):
        c_arr = np.asarray(c.detach().cpu() if isinstance(c, torch.Tensor) else c, dtype=np.float64).reshape(-1)
        dim = c_arr.size if dim is None else dim
        super().__init__(dim, scale)
        self.c = _vec(c, self.dim, 1.0)
        self.w = float(w)

    def _fill(self, d):
        d.c1 = self.w
        for k in range(self.dim):
            d.a[k] = self.c[k]

    def torch_eval(self, x):
        c = self._t(self.c, x)
        return torch.cos(2.0 * math.pi * self.w + x @ c)


class ProductPeak(BuiltinIntegrand):
    family = _lib.FAM_PRODUCT_PEAK
    name = "ProductPeak"

    def __init__(self, a, u, dim=None, scale=1.0):
        a_arr = np.asarray(a.detach().cpu() if isinstance(a, torch.Tensor) else a, dtype=np.float64).reshape(-1)
        dim = a_arr.size if dim is None else dim
        super().__init__(dim, scale)
        self.a = _vec(a, self.dim, 1.0)
        self.u = _vec(u, self.dim, 0.0)
        if np.any(self.a == 0.0):
            raise ValueError("ProductPeak needs a_k != 0")

    def _fill(self, d):
        for k in range(self.dim):
            d.a[k] = self.a[k]
            d.u[k] = self.u[k]

    def torch_eval(self, x):
        a, u = self._t(self.a, x), self._t(self.u, x)
        return torch.prod(1.0 / (a ** -2 + (x - u) ** 2), dim=-1)


class ExpLinear(BuiltinIntegrand):
    family = _lib.FAM_EXP_LINEAR
    name = "ExpLinear"

    def __init__(self, dim, a=None, c=0.0, scale=1.0):
        super().__init__(dim, scale)
        self.a = _vec(a, self.dim, 1.0)
        self.c = float(c)

    def _fill(self, d):
        d.c1 = self.c
        for k in range(self.dim):
            d.a[k] = self.a[k]

    def torch_eval(self, x):
        return torch.exp(self.c + x @ self._t(self.a, x))


class SinSum(BuiltinIntegrand):
    family = _lib.FAM_SIN_SUM
    name = "SinSum"

    def __init__(self, dim, a=None, u=None, scale=1.0):
        super().__init__(dim, scale)
        self.a = _vec(a, self.dim, 1.0)
        self.u = _vec(u, self.dim, 0.0)

    def _fill(self, d):
        for k in range(self.dim):
            d.a[k] = self.a[k]
            d.u[k] = self.u[k]

    def torch_eval(self, x):
        return torch.sum(torch.sin(x * self._t(self.a, x) + self._t(self.u, x)), dim=-1)


class DoubleGaussian(BuiltinIntegrand):
    """Two Gaussian peaks on the diagonal, normalised to integrate to ~1."""

    family = _lib.FAM_DOUBLE_GAUSSIAN
    name = "DoubleGaussian"

    def __init__(self, dim, width=0.1, centers=(1.0 / 3.0, 2.0 / 3.0), scale=1.0):
        super().__init__(dim, scale)
        self.width = float(width)
        self.centers = (float(centers[0]), float(centers[1]))
        self.norm = 0.5 * (self.width * math.sqrt(math.pi)) ** (-self.dim)

    def descriptor(self, dim=None):
        d = super().descriptor(dim)
        d.c0 = self.scale * self.norm
        return d

    def _fill(self, d):
        d.c1, d.c2 = self.centers
        d.c3 = 1.0 / (self.width * self.width)

    def torch_eval(self, x):
        inv = 1.0 / (self.width * self.width)
        s1 = torch.sum((x - self.centers[0]) ** 2, dim=-1)
        s2 = torch.sum((x - self.centers[1]) ** 2, dim=-1)
        return self.norm * (torch.exp(-inv * s1) + torch.exp(-inv * s2))


class Constant(BuiltinIntegrand):
    family = _lib.FAM_CONSTANT
    name = "Constant"

    def __init__(self, dim, value=1.0):
        super().__init__(dim, scale=value)

    def _fill(self, d):
        pass

    def torch_eval(self, x):
        return torch.ones(x.shape[:-1], dtype=x.dtype, device=x.device)


def is_builtin(fn, dim=None):
    """True when `fn` is a built-in family usable by the fused kernels."""
    if not isinstance(fn, BuiltinIntegrand):
        return False
    if dim is not None and fn.dim != dim:
        return False
    return fn.dim <= _lib.TQ_MAX_DIM


__all__ = [
    "BuiltinIntegrand",
    "GaussianPeak",
    "Oscillatory",
    "ProductPeak",
    "ExpLinear",
    "SinSum",
    "DoubleGaussian",
    "Constant",
    "is_builtin",
]
