"""Domain handling and helpers shared by the integrators.

Mirrors torchquad/integration/utils.py (the torch backend):
``_linspace_with_grads`` (:24-58), ``_setup_integration_domain`` (:110-159),
``_check_integration_domain`` (:162-206) and
``expand_func_values_and_squeeze_integral`` (:235-277).  Error messages and
warnings are the reference's.
"""
import warnings
import weakref

import torch

from ..utils.set_log_level import logger
from ..utils.set_precision import _get_precision
from ..utils.set_up_backend import _get_default_backend

_SUPPORTED_BACKENDS = ("torch",)


def _default_device():
    """GPU when one is visible (this package computes only there), else CPU."""
    dev = torch.get_default_device() if hasattr(torch, "get_default_device") else torch.device("cpu")
    if dev.type == "cuda":
        return dev
    if torch.cuda.is_available():
        return torch.device("cuda", torch.cuda.current_device())
    return dev


def _linspace_with_grads(start, stop, N, requires_grad):
    """Equally spaced 1-D grid that keeps gradients w.r.t. start/stop (utils.py:24-58)."""
    if requires_grad:
        zero = start.new_zeros(())
        one = start.new_ones(())
        grid = torch.linspace(zero, one, N, dtype=start.dtype, device=start.device)
        grid = grid * (stop - start)
        grid = grid + start
        return grid
    return torch.linspace(start, stop, N, dtype=start.dtype, device=start.device)


def _setup_integration_domain(dim, integration_domain, backend):
    """List/None -> tensor [dim, 2]; validate the shape (utils.py:110-159).

    A list becomes a tensor on the GPU with dtype from TORCHQUAD_DTYPE_TORCH
    or torch's default dtype.
    """
    if backend is not None and backend not in _SUPPORTED_BACKENDS:
        raise ValueError(
            f"Unsupported numerical backend: {backend}. torchquad_amd computes with backend='torch'."
        )
    if integration_domain is None:
        integration_domain = [[-1.0, 1.0]] * dim

    is_tensor = isinstance(integration_domain, torch.Tensor)
    if not is_tensor and hasattr(integration_domain, "__array__") and not isinstance(
        integration_domain, (list, tuple)
    ):
        # numpy arrays etc.: the reference keeps their backend; here they are
        # converted like a list when the torch backend is requested.
        if backend is None and _get_default_backend() not in _SUPPORTED_BACKENDS:
            raise ValueError("torchquad_amd computes with backend='torch' only")
        warning_msg = "integration_domain should be a list when the backend argument is set."
        logger.warning(warning_msg)
        warnings.warn(warning_msg, RuntimeWarning)
        integration_domain = integration_domain.tolist()

    if not isinstance(integration_domain, torch.Tensor):
        integration_domain = [[float(b) for b in bounds] for bounds in integration_domain]
        dtype_name = _get_precision("torch")
        dtype = {"float32": torch.float32, "float64": torch.float64}.get(dtype_name, torch.get_default_dtype())
        _check_list_domain(integration_domain)
        host = integration_domain
        integration_domain = torch.tensor(integration_domain, dtype=dtype, device=_default_device())
        _mark_checked(integration_domain)
        # the values as the tensor holds them (rounded to its dtype)
        _remember_host(integration_domain, torch.tensor(host, dtype=dtype).to(torch.float64).numpy())

    if tuple(integration_domain.shape) != (dim, 2):
        raise ValueError(
            "The integration domain has an unexpected shape. "
            f"Expected {(dim, 2)}, got {tuple(integration_domain.shape)}"
        )
    return integration_domain


def _check_list_domain(integration_domain):
    dim = len(integration_domain)
    if dim < 1:
        raise ValueError("len(integration_domain) needs to be 1 or larger.")
    for bounds in integration_domain:
        if len(bounds) != 2:
            raise ValueError(bounds, " in ", integration_domain, " does not specify a valid integration bound.")
        if bounds[0] > bounds[1]:
            raise ValueError(bounds, " in ", integration_domain, " does not specify a valid integration bound.")
    return dim


# Tensors whose bounds were validated, keyed by identity and in-place
# version counter: validation of a device tensor needs one device->host read,
# which is paid once per tensor instead of once per call.
_checked = {}


def _mark_checked(t):
    key = id(t)
    try:
        ref = weakref.ref(t, lambda _r, k=key: _checked.pop(k, None))
    except TypeError:
        return
    _checked[key] = (ref, t.data_ptr(), t._version)


def _is_marked(t):
    entry = _checked.get(id(t))
    if entry is None:
        return False
    ref, p, v = entry
    return ref() is t and p == t.data_ptr() and v == t._version


_ADD_AT_DTYPES = {torch.float32: 0, torch.float64: 1, torch.int32: 2, torch.int64: 3}


def _add_at_indices(target, indices, source, is_sorted=False):
    """Add source[i] to target[indices[i]] for every i, in place (utils.py:61-107).

    For example target=[0,0,0], indices=[2,1,1,2], source=[a,b,c,d] gives
    [0, b+c, a+d].  target, indices and source are 1-D tensors on the GPU;
    target and source share a dtype (float32, float64, int32 or int64) and
    indices are integers.  One HIP kernel (tq_add_at_indices) does the work;
    is_sorted is accepted for the reference's signature and needs no care:
    each wave checks whether its indices ascend and sorts them if not.
    Integer sums are exact; float sums depend on the order in which waves
    reach a shared entry, as torch's scatter_add_ on a GPU does.

    Raises:
        ValueError: shapes, dtypes or devices that do not fit.
        IndexError: an index outside [0, len(target)) (checked after the
            kernel, which skips such entries; this synchronises the stream).
    """
    from .. import _lib

    for name, t in (("target", target), ("indices", indices), ("source", source)):
        if not isinstance(t, torch.Tensor):
            raise ValueError(f"{name} must be a torch tensor")
        _lib.require_device(t)
    if target.dim() != 1 or indices.dim() != 1 or source.shape != indices.shape:
        raise ValueError("expected a 1-D target and 1-D indices and source of one length")
    if target.dtype != source.dtype or target.dtype not in _ADD_AT_DTYPES:
        raise ValueError("target and source must share a dtype among float32, float64, int32 and int64")
    if indices.dtype.is_floating_point or indices.dtype == torch.bool:
        raise ValueError("indices must be an integer tensor")
    if not (target.device == indices.device == source.device):
        raise ValueError("target, indices and source must be on one device")
    if indices.numel() == 0:
        return
    idx = indices.to(torch.int64).contiguous()
    src = source.contiguous()
    dst = target if target.is_contiguous() else target.contiguous()
    status = torch.zeros(1, dtype=torch.int32, device=target.device)
    _lib.call("tq_add_at_indices", _ADD_AT_DTYPES[target.dtype], _lib.ptr(dst), dst.numel(), _lib.ptr(idx),
              _lib.ptr(src), idx.numel(), _lib.ptr(status), _lib.stream_ptr(target.device))
    if dst is not target:
        target.copy_(dst)
    if int(status.item()):
        raise IndexError(f"_add_at_indices: an index is outside [0, {target.numel()})")


def _check_integration_domain(integration_domain):
    """Validate shape and bounds and return the dimension (utils.py:162-206)."""
    if not isinstance(integration_domain, torch.Tensor):
        return _check_list_domain(integration_domain)
    if len(integration_domain.shape) != 2:
        raise ValueError("The integration_domain tensor has an invalid shape")
    dim, num_bounds = integration_domain.shape
    if dim < 1:
        raise ValueError("integration_domain.shape[0] needs to be 1 or larger.")
    if num_bounds != 2:
        raise ValueError("integration_domain must have 2 values per boundary")
    if torch.jit.is_tracing():
        return dim
    if _is_marked(integration_domain):
        return dim
    if torch.min(integration_domain[:, 1] - integration_domain[:, 0]).item() < 0.0:
        raise ValueError("integration_domain has invalid boundary values")
    _mark_checked(integration_domain)
    return dim


# Host copies of device domains, keyed like _checked: the batched kernels
# take domains by value, and a domain tensor reused across calls is read to
# the host once.
_host_copies = {}


def _remember_host(t, values):
    key = id(t)
    try:
        ref = weakref.ref(t, lambda _r, k=key: _host_copies.pop(k, None))
    except TypeError:
        return
    _host_copies[key] = (ref, t.data_ptr(), t._version, values)


def _host_values(t):
    """float64 numpy copy of a domain tensor (one device read per tensor and
    in-place version)."""
    e = _host_copies.get(id(t))
    if e is not None:
        ref, p, v, values = e
        if ref() is t and p == t.data_ptr() and v == t._version:
            return values
    values = t.detach().to("cpu", torch.float64).numpy().copy()
    _remember_host(t, values)
    return values


def _float_domain(integration_domain):
    """Integer domains become float64 (integration_grid.py:52-56); contiguous."""
    if not integration_domain.dtype.is_floating_point:
        integration_domain = integration_domain.to(torch.float64)
    return integration_domain.contiguous()


def expand_func_values_and_squeeze_integral(f):
    """Decorator: [N] values become [N, 1] and the result is squeezed, with the
    reference's deprecation warning (utils.py:235-277)."""

    def wrap(*args, **kwargs):
        if len(args) > 1:
            function_values = args[1]
        elif "function_values" in kwargs:
            function_values = kwargs["function_values"]
        else:
            raise ValueError(
                "function_values argument not found in either positional or keyword arguments. "
                "Please provide function_values as the second positional argument or as a keyword argument."
            )
        is_1d = len(function_values.shape) == 1 or (
            len(function_values.shape) == 2 and function_values.shape[1] == 1
        )
        if is_1d:
            warnings.warn(
                "DEPRECATION WARNING: In future versions of torchquad, an array-like object will be returned."
            )
            if len(args) > 1:
                args = (args[0], torch.unsqueeze(function_values, 1) if function_values.dim() == 1
                        else function_values, *args[2:])
            else:
                fv = kwargs["function_values"]
                kwargs["function_values"] = torch.unsqueeze(fv, 1) if fv.dim() == 1 else fv
            result = f(*args, **kwargs)
            return torch.squeeze(result)
        return f(*args, **kwargs)

    return wrap


def _warn_scalar_deprecation():
    """The warning the reference emits for every scalar integrand (utils.py:264-266)."""
    warnings.warn(
        "DEPRECATION WARNING: In future versions of torchquad, an array-like object will be returned."
    )
