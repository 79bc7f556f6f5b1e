"""Built-in integrand families that the fused HIP kernels evaluate in-register.

Passing one of these objects as ``fn`` to an integrator selects the fused
path: points are never written to memory.  Each object is also an ordinary
callable on a ``[N, dim]`` tensor (its mathematical definition in torch), so
it can be handed to any code that expects a torchquad integrand, and the
generic user-integrand path can be compared against the fused one.

Families (per-axis parameters are vectors of length ``dim``):

* ``GaussianPeak(a, u)``      f = exp(-sum a_k^2 (x_k - u_k)^2)       (Genz "Gaussian")
* ``Oscillatory(c, w)``       f = cos(2 pi w + sum c_k x_k)           (Genz "oscillatory")
* ``ProductPeak(a, u)``       f = prod 1 / (a_k^-2 + (x_k - u_k)^2)   (Genz "product peak")
* ``ExpLinear(a, c)``         f = exp(c + sum a_k x_k)
* ``SinSum(a, u)``            f = sum sin(a_k x_k + u_k)
* ``DoubleGaussian(width, centers)``
  f = 1/2 sum_{c in centers} (width sqrt(pi))^-dim exp(-sum (x_k - c)^2 / width^2)
* ``Constant(value)``         f = value

Parameter batches.  Any parameter may carry leading batch dimensions S:
per-axis parameters of shape ``[*S, dim]``, scalar ones of shape ``[*S]``
(all broadcast together).  Such an object is the vectorised integrand of the
reference (base_integrator.py:77-89, tutorial.rst:741-774): called on
``[N, dim]`` points it returns ``[N, *S]``, and an integrator returns the
``[*S]`` integrals.  On the fused path each of the prod(S) parameter sets is
one integral of a single batched launch over the same grid.
"""
import math

import numpy as np
import torch

from . import _lib


def _array(v, default):
    if v is None:
        return np.asarray(float(default), dtype=np.float64)
    if isinstance(v, torch.Tensor):
        v = v.detach().to("cpu", torch.float64).numpy()
    return np.asarray(v, dtype=np.float64)


def _axis_param(v, dim, default, name):
    """[*S, dim] array of a per-axis parameter (a scalar or a length-1 last
    axis is repeated over the axes)."""
    arr = _array(v, default)
    if arr.ndim == 0:
        return np.full(dim, float(arr))
    if arr.shape[-1] == 1 and dim > 1:
        arr = np.repeat(arr, dim, axis=-1)
    if arr.shape[-1] != dim:
        raise ValueError(f"parameter {name} has {arr.shape[-1]} entries per set, expected dim={dim}")
    return arr


def _dim_of(v):
    arr = _array(v, 1.0)
    return 1 if arr.ndim == 0 else int(arr.shape[-1])


class BuiltinIntegrand:
    """Base class: a family id, a dimension and float64 parameters.

    Subclasses register their parameters with ``_set_params``: per-axis ones
    (``[*S, dim]``) and scalar ones (``[*S]``); ``batch_shape`` is the
    broadcast S (``()`` for a single integrand).
    """

    family = None
    name = "builtin"

    def __init__(self, dim, scale=1.0):
        self.dim = int(dim)
        if self.dim < 1:
            raise ValueError("dim must be >= 1")
        self._axis_names = ()
        self._scalar_names = ("scale",)
        self.scale = _array(scale, 1.0)
        self.batch_shape = tuple(self.scale.shape)

    def _set_params(self, axis=None, scalar=None):
        axis = dict(axis or {})
        scalar = dict(scalar or {})
        scalar.setdefault("scale", self.scale)
        shapes = [np.shape(v)[:-1] for v in axis.values()] + [np.shape(v) for v in scalar.values()]
        try:
            S = tuple(np.broadcast_shapes(*shapes)) if shapes else ()
        except ValueError as e:
            raise ValueError(f"{self.name}: parameter batch shapes {shapes} do not broadcast") from e
        for k, v in axis.items():
            setattr(self, k, np.array(np.broadcast_to(v, S + (self.dim,))))
        for k, v in scalar.items():
            v = np.broadcast_to(v, S)
            setattr(self, k, np.array(v) if S else float(v))
        self._axis_names = tuple(axis)
        self._scalar_names = tuple(scalar)
        self.batch_shape = S

    @property
    def batch_size(self):
        return int(np.prod(self.batch_shape)) if self.batch_shape else 1

    # -- fused-kernel descriptors ------------------------------------------
    def _fill(self, d, i):  # pragma: no cover - overridden
        raise NotImplementedError

    def _ax(self, name, i):
        v = getattr(self, name)
        return v if not self.batch_shape else v.reshape(-1, self.dim)[i]

    def _sc(self, name, i):
        v = getattr(self, name)
        return float(v) if not self.batch_shape else float(np.reshape(v, -1)[i])

    def _check_dim(self, dim):
        dim = self.dim if dim is None else int(dim)
        if dim != self.dim:
            raise ValueError(f"{self.name} was built for dim={self.dim}, called with dim={dim}")
        if dim > _lib.TQ_MAX_DIM:
            raise ValueError(f"fused integrands support dim <= {_lib.TQ_MAX_DIM}")
        return dim

    def _descriptor(self, dim, i):
        d = _lib.Integrand()
        d.family = self.family
        d.dim = dim
        d.c0 = self._sc("scale", i)
        self._fill(d, i)
        return d

    def descriptor(self, dim=None):
        """The C descriptor of a single integrand (no parameter batch)."""
        dim = self._check_dim(dim)
        if self.batch_shape:
            raise ValueError(f"{self.name} holds a parameter batch of shape {self.batch_shape}; use descriptors()")
        return self._descriptor(dim, 0)

    def descriptors(self, dim=None):
        """C descriptors of every parameter set, in C order of the batch shape."""
        dim = self._check_dim(dim)
        return [self._descriptor(dim, i) for i in range(self.batch_size)]

    def key(self):
        """Bytes identifying the family, dimension and every parameter value."""
        parts = [str((self.family, self.dim, self.batch_shape)).encode()]
        for name in self._axis_names + self._scalar_names:
            parts.append(np.ascontiguousarray(getattr(self, name), dtype=np.float64).tobytes())
        return b"|".join(parts)

    # -- torch definition -------------------------------------------------
    def torch_eval(self, x):  # pragma: no cover - overridden
        raise NotImplementedError

    def __call__(self, x):
        if x.shape[-1] != self.dim:
            raise ValueError(f"{self.name} expects points of dim {self.dim}, got {x.shape[-1]}")
        if self.batch_shape:
            # [N, dim] -> [N, 1, ..., 1, dim] against parameters [*S, dim]: [N, *S]
            x = x.reshape((x.shape[0],) + (1,) * len(self.batch_shape) + (self.dim,))
        return self._t(self.scale, x) * self.torch_eval(x)

    def _t(self, arr, x):
        return torch.as_tensor(arr, dtype=x.dtype, device=x.device)

    def __repr__(self):
        s = f", batch_shape={self.batch_shape}" if self.batch_shape else ""
        return f"{type(self).__name__}(dim={self.dim}{s})"


class GaussianPeak(BuiltinIntegrand):
    family = _lib.FAM_GAUSSIAN_PEAK
    name = "GaussianPeak"

    def __init__(self, a, u, dim=None, scale=1.0):
        super().__init__(_dim_of(a) if dim is None else dim, scale)
        self._set_params(axis={"a": _axis_param(a, self.dim, 1.0, "a"), "u": _axis_param(u, self.dim, 0.0, "u")})

    def _fill(self, d, i):
        a, u = self._ax("a", i), self._ax("u", i)
        for k in range(self.dim):
            d.a[k] = a[k]
            d.u[k] = u[k]

    def torch_eval(self, x):
        a, u = self._t(self.a, x), self._t(self.u, x)
        return torch.exp(-torch.sum((a * (x - u)) ** 2, dim=-1))


class Oscillatory(BuiltinIntegrand):
    family = _lib.FAM_OSCILLATORY
    name = "Oscillatory"

    def __init__(self, c, w=0.0, dim=None, scale=1.0):
        super().__init__(_dim_of(c) if dim is None else dim, scale)
        self._set_params(axis={"c": _axis_param(c, self.dim, 1.0, "c")}, scalar={"w": _array(w, 0.0)})

    def _fill(self, d, i):
        d.c1 = self._sc("w", i)
        c = self._ax("c", i)
        for k in range(self.dim):
            d.a[k] = c[k]

    def torch_eval(self, x):
        c = self._t(self.c, x)
        return torch.cos(2.0 * math.pi * self._t(self.w, x) + torch.sum(x * c, dim=-1))


class ProductPeak(BuiltinIntegrand):
    family = _lib.FAM_PRODUCT_PEAK
    name = "ProductPeak"

    def __init__(self, a, u, dim=None, scale=1.0):
        super().__init__(_dim_of(a) if dim is None else dim, scale)
        self._set_params(axis={"a": _axis_param(a, self.dim, 1.0, "a"), "u": _axis_param(u, self.dim, 0.0, "u")})
        if np.any(self.a == 0.0):
            raise ValueError("ProductPeak needs a_k != 0")

    def _fill(self, d, i):
        a, u = self._ax("a", i), self._ax("u", i)
        for k in range(self.dim):
            d.a[k] = a[k]
            d.u[k] = u[k]

    def torch_eval(self, x):
        a, u = self._t(self.a, x), self._t(self.u, x)
        return torch.prod(1.0 / (a ** -2 + (x - u) ** 2), dim=-1)


class ExpLinear(BuiltinIntegrand):
    family = _lib.FAM_EXP_LINEAR
    name = "ExpLinear"

    def __init__(self, dim, a=None, c=0.0, scale=1.0):
        super().__init__(dim, scale)
        self._set_params(axis={"a": _axis_param(a, self.dim, 1.0, "a")}, scalar={"c": _array(c, 0.0)})

    def _fill(self, d, i):
        d.c1 = self._sc("c", i)
        a = self._ax("a", i)
        for k in range(self.dim):
            d.a[k] = a[k]

    def torch_eval(self, x):
        return torch.exp(self._t(self.c, x) + torch.sum(x * self._t(self.a, x), dim=-1))


class SinSum(BuiltinIntegrand):
    family = _lib.FAM_SIN_SUM
    name = "SinSum"

    def __init__(self, dim, a=None, u=None, scale=1.0):
        super().__init__(dim, scale)
        self._set_params(axis={"a": _axis_param(a, self.dim, 1.0, "a"), "u": _axis_param(u, self.dim, 0.0, "u")})

    def _fill(self, d, i):
        a, u = self._ax("a", i), self._ax("u", i)
        for k in range(self.dim):
            d.a[k] = a[k]
            d.u[k] = u[k]

    def torch_eval(self, x):
        return torch.sum(torch.sin(x * self._t(self.a, x) + self._t(self.u, x)), dim=-1)


class DoubleGaussian(BuiltinIntegrand):
    """Two Gaussian peaks on the diagonal, normalised to integrate to ~1."""

    family = _lib.FAM_DOUBLE_GAUSSIAN
    name = "DoubleGaussian"

    def __init__(self, dim, width=0.1, centers=(1.0 / 3.0, 2.0 / 3.0), scale=1.0):
        super().__init__(dim, scale)
        self.centers = (float(centers[0]), float(centers[1]))
        self._set_params(scalar={"width": _array(width, 0.1)})
        self.norm = 0.5 * (self.width * math.sqrt(math.pi)) ** (-self.dim)

    def _fill(self, d, i):
        w = self._sc("width", i)
        d.c0 = self._sc("scale", i) * 0.5 * (w * math.sqrt(math.pi)) ** (-self.dim)
        d.c1, d.c2 = self.centers
        d.c3 = 1.0 / (w * w)

    def torch_eval(self, x):
        w = self._t(self.width, x)
        inv = 1.0 / (w * w)
        s1 = torch.sum((x - self.centers[0]) ** 2, dim=-1)
        s2 = torch.sum((x - self.centers[1]) ** 2, dim=-1)
        return self._t(self.norm, x) * (torch.exp(-inv * s1) + torch.exp(-inv * s2))


class Constant(BuiltinIntegrand):
    family = _lib.FAM_CONSTANT
    name = "Constant"

    def __init__(self, dim, value=1.0):
        super().__init__(dim, scale=value)
        self._set_params()

    def _fill(self, d, i):
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


def is_batched(fn):
    """True when `fn` is a built-in family holding a parameter batch."""
    return isinstance(fn, BuiltinIntegrand) and bool(fn.batch_shape)


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
    "is_batched",
]
