"""ctypes binding of ``libtorchquad_amd.so`` (declared in include/torchquad_amd.h).

The library is the only compute path of this package: every integrator
launches its kernels through the functions bound here.  If the library is
missing or no ROCm device is visible, compute calls raise
:class:`NativeUnavailableError` instead of computing anything elsewhere.
"""
import ctypes
import os
import threading
from pathlib import Path

import torch

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libtorchquad_amd.so"

TQ_MAX_DIM = 16
TQ_F32, TQ_F64, TQ_I32, TQ_I64 = 0, 1, 2, 3
RULE_TRAPEZOID, RULE_SIMPSON, RULE_BOOLE, RULE_GAUSS_LEGENDRE = 0, 1, 2, 3

FAM_GAUSSIAN_PEAK = 0
FAM_OSCILLATORY = 1
FAM_PRODUCT_PEAK = 2
FAM_EXP_LINEAR = 3
FAM_SIN_SUM = 4
FAM_DOUBLE_GAUSSIAN = 5
FAM_CONSTANT = 6


class NativeUnavailableError(RuntimeError):
    """The HIP library or a ROCm GPU is not available for a compute call."""


class NativeError(RuntimeError):
    """A C ABI call returned a non-zero status."""

    def __init__(self, fn, code, msg):
        super().__init__(f"{fn} failed with status {code}: {msg}")
        self.code = code


class Integrand(ctypes.Structure):
    _fields_ = [
        ("family", ctypes.c_int32),
        ("dim", ctypes.c_int32),
        ("c0", ctypes.c_double),
        ("c1", ctypes.c_double),
        ("c2", ctypes.c_double),
        ("c3", ctypes.c_double),
        ("a", ctypes.c_double * TQ_MAX_DIM),
        ("u", ctypes.c_double * TQ_MAX_DIM),
    ]


class GridDesc(ctypes.Structure):
    _fields_ = [
        ("dtype", ctypes.c_int32),
        ("rule", ctypes.c_int32),
        ("dim", ctypes.c_int32),
        ("n", ctypes.c_int32),
        ("lane_axes", ctypes.c_int32),
        ("items_per_wave", ctypes.c_int32),
        ("gl_nodes", ctypes.c_void_p),
        ("gl_weights", ctypes.c_void_p),
    ]


class BatchItem(ctypes.Structure):
    _fields_ = [
        ("rule", ctypes.c_int32),
        ("family", ctypes.c_int32),
        ("dim", ctypes.c_int32),
        ("n", ctypes.c_int32),
        ("gl_offset", ctypes.c_int64),
        ("domain", ctypes.c_double * (2 * TQ_MAX_DIM)),
        ("c0", ctypes.c_double),
        ("c1", ctypes.c_double),
        ("c2", ctypes.c_double),
        ("c3", ctypes.c_double),
        ("a", ctypes.c_double * TQ_MAX_DIM),
        ("u", ctypes.c_double * TQ_MAX_DIM),
    ]


class BatchTile(ctypes.Structure):
    _fields_ = [
        ("item", ctypes.c_int32),
        ("lane_wave", ctypes.c_int32),
        ("lane_wave_end", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("t0", ctypes.c_int64),
        ("t1", ctypes.c_int64),
    ]


class McDesc(ctypes.Structure):
    _fields_ = [
        ("dtype", ctypes.c_int32),
        ("dim", ctypes.c_int32),
        ("N", ctypes.c_int64),
        ("seed", ctypes.c_uint64),
        ("offset", ctypes.c_uint64),
        ("samples_per_thread", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


_P = ctypes.c_void_p
_SZ = ctypes.c_size_t
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_U64 = ctypes.c_uint64

# name -> (restype, argtypes)
SIGNATURES = {
    "tq_last_error": (ctypes.c_char_p, []),
    "tq_version": (ctypes.c_int, []),
    "tq_device_count": (ctypes.c_int, []),
    "tq_grid_workspace_size": (_SZ, [ctypes.POINTER(GridDesc)]),
    "tq_grid_integrate": (
        ctypes.c_int,
        [ctypes.POINTER(GridDesc), ctypes.POINTER(Integrand), _P, _P, _SZ, _P, _P, _P],
    ),
    "tq_grid_integrate_noalloc": (
        ctypes.c_int,
        [ctypes.POINTER(GridDesc), ctypes.POINTER(Integrand), _P, _P, _P, _P],
    ),
    "tq_grid_points": (ctypes.c_int, [ctypes.POINTER(GridDesc), _P, _P, _P]),
    "tq_grid_h": (ctypes.c_int, [ctypes.POINTER(GridDesc), _P, _P, _P]),
    "tq_grid_reduce_workspace_size": (_SZ, [ctypes.POINTER(GridDesc), _I64]),
    "tq_grid_reduce": (ctypes.c_int, [ctypes.POINTER(GridDesc), _P, _P, _I32, _P, _I64, _P, _SZ, _P, _P]),
    "tq_grid_reduce_backward": (ctypes.c_int, [ctypes.POINTER(GridDesc), _P, _P, _I32, _P, _I64, _P, _P]),
    "tq_grid_points_backward_workspace_size": (_SZ, [ctypes.POINTER(GridDesc)]),
    "tq_grid_points_backward": (ctypes.c_int, [ctypes.POINTER(GridDesc), _P, _P, _P, _SZ, _P, _P]),
    "tq_gauss_legendre": (ctypes.c_int, [_I32, _P, _P, _P]),
    "tq_sweep_workspace_size": (_SZ, [ctypes.POINTER(GridDesc), _I64]),
    "tq_sweep_gaussian_forward": (ctypes.c_int, [ctypes.POINTER(GridDesc), _I64, _P, _P, _P, _P, _SZ, _P, _P, _P]),
    "tq_sweep_gaussian_backward": (
        ctypes.c_int,
        [ctypes.POINTER(GridDesc), _I64, _P, _P, _P, _P, _P, _SZ, _P, _P, _P, _P],
    ),
    "tq_probe_rates": (ctypes.c_int, [_I32, _I64, _I32, _P, _P]),
    "tq_sweep_forward": (ctypes.c_int, [ctypes.POINTER(GridDesc), _I32, _I64, _P, _P, _P, _P, _SZ, _P, _P, _P]),
    "tq_sweep_backward": (
        ctypes.c_int,
        [ctypes.POINTER(GridDesc), _I32, _I64, _P, _P, _P, _P, _P, _SZ, _P, _P, _P, _P],
    ),
    "tq_batch_plan_sizes": (ctypes.c_int, [_P, _I64, _I32, _I64, ctypes.POINTER(_I64), ctypes.POINTER(_I64)]),
    "tq_batch_plan": (ctypes.c_int, [_P, _I64, _I32, _I64, _P, _P, _P, _P, _P]),
    "tq_batch_run": (ctypes.c_int, [_I32, _P, _I64, _P, _I64, _P, _P, _P, _P, _P, _I64, _P, _P, _P, _P, _P]),
    "tq_mc_workspace_size": (_SZ, [ctypes.POINTER(McDesc)]),
    "tq_mc_integrate": (
        ctypes.c_int,
        [ctypes.POINTER(McDesc), ctypes.POINTER(Integrand), _P, _P, _SZ, _P, _P, _P],
    ),
    "tq_mc_integrate_dev_offset": (
        ctypes.c_int,
        [ctypes.POINTER(McDesc), ctypes.POINTER(Integrand), _P, _P, _P, _SZ, _P, _P, _P],
    ),
    "tq_mc_points": (ctypes.c_int, [ctypes.POINTER(McDesc), _P, _P, _P]),
    "tq_mc_reduce_workspace_size": (_SZ, [_I64, _I64]),
    "tq_mc_reduce": (ctypes.c_int, [_I32, _P, _I64, _I64, _P, _I32, _P, _SZ, _P, _P]),
    "tq_uniform": (ctypes.c_int, [_U64, _U64, _I32, _I64, _I32, _P, _P]),
}

class VegasState(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int32),
        ("n_intervals", ctypes.c_int32),
        ("n_strat", ctypes.c_int32),
        ("sample_dtype", ctypes.c_int32),
        ("n_cubes", ctypes.c_int64),
        ("max_samples", ctypes.c_int64),
        ("max_blocks", ctypes.c_int64),
        ("v_cubes", ctypes.c_double),
        ("alpha", ctypes.c_double),
        ("beta", ctypes.c_double),
        ("x_edges", ctypes.c_void_p),
        ("dx_edges", ctypes.c_void_p),
        ("weights", ctypes.c_void_p),
        ("counts", ctypes.c_void_p),
        ("dh", ctypes.c_void_p),
        ("jf_cube", ctypes.c_void_p),
        ("jf2_cube", ctypes.c_void_p),
        ("nh", ctypes.c_void_p),
        ("offsets", ctypes.c_void_p),
        ("chunk_cube", ctypes.c_void_p),
        ("jf", ctypes.c_void_p),
        ("hist_w", ctypes.c_void_p),
        ("hist_c", ctypes.c_void_p),
        ("block_sums", ctypes.c_void_p),
        ("scan_sums", ctypes.c_void_p),
        ("iter_stats", ctypes.c_void_p),
        ("flags", ctypes.c_void_p),
    ]


_P = ctypes.c_void_p
_I32, _I64, _U64, _D = ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64, ctypes.c_double
_S = ctypes.POINTER(VegasState)
VEGAS_SIGNATURES = {
    "tq_vegas_init": (ctypes.c_int, [_S, _P]),
    "tq_vegas_nh": (ctypes.c_int, [_S, _I64, _I32, _I32, _P, _P]),
    "tq_vegas_iterate": (
        ctypes.c_int,
        [_S, ctypes.POINTER(Integrand), _P, _U64, _U64, _I32, _I32, _I32, _I64, _P],
    ),
    "tq_vegas_points": (ctypes.c_int, [_S, _P, _U64, _U64, _I32, _I64, _P, _P, _P, _P, _P, _P]),
    "tq_vegas_consume": (ctypes.c_int, [_S, _P, _P, _P, _P, _I32, _I32, _I32, _I64, _P]),
    "tq_vegas_update_map": (ctypes.c_int, [_S, _I64, _I32, _P]),
    "tq_vegas_map_eval": (ctypes.c_int, [_S, _P, _I64, _P, _P, _P, _P, _P]),
    "tq_vegas_map_accumulate": (ctypes.c_int, [_S, _P, _P, _I64, _P]),
    "tq_vegas_map_update_accumulated": (ctypes.c_int, [_S, _P]),
    "tq_vegas_smooth": (ctypes.c_int, [_I32, _I32, _P, _P, _D, _P, _P, _P, _P]),
    "tq_vegas_strat_y": (ctypes.c_int, [_S, _U64, _U64, _P, _P, _P]),
    "tq_vegas_strat_accumulate": (ctypes.c_int, [_S, _P, _I32, _P]),
    "tq_vegas_update_dh": (ctypes.c_int, [_S, _I32, _P]),
    "tq_vegas_normalize_dh": (ctypes.c_int, [_S, _I32, _P]),
    "tq_add_at_indices": (ctypes.c_int, [_I32, _P, _I64, _P, _P, _I64, _P, _P]),
}


SIGNATURES.update(VEGAS_SIGNATURES)

_lib = None
_lib_lock = threading.Lock()


def load_library(path=None):
    """Load and bind the shared library (no GPU needed to load it)."""
    global _lib
    if _lib is not None and path is None:  # bound once; read without the lock
        return _lib
    with _lib_lock:
        if _lib is not None and path is None:
            return _lib
        p = Path(path) if path else LIB_PATH
        if not p.exists():
            raise NativeUnavailableError(
                f"{p} is missing; build it with `python -m torchquad_amd._build`"
            )
        lib = ctypes.CDLL(str(p))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        if path is None:
            _lib = lib
        return lib


def lib():
    return load_library()


_device_ready = False  # a GPU was seen and the library bound (checked once)


def require_device(tensor):
    """Raise unless `tensor` lives on a ROCm GPU that the library can use."""
    global _device_ready
    if not _device_ready and not torch.cuda.is_available():
        raise NativeUnavailableError(
            "torchquad_amd computes on a ROCm GPU only and none is visible to torch"
        )
    if tensor.device.type != "cuda":
        raise NativeUnavailableError(
            "torchquad_amd computes on a ROCm GPU only; the integration domain is on "
            f"{tensor.device}. Call enable_cuda() / set_up_backend('torch') or pass a "
            "domain tensor on the GPU."
        )
    if not _device_ready:
        lib()
        _device_ready = True


def call(name, *args):
    """Call a status-returning entry point and raise NativeError on failure."""
    fn = getattr(lib(), name)
    rc = fn(*args)
    if rc != 0:
        msg = lib().tq_last_error().decode(errors="replace")
        raise NativeError(name, rc, msg)
    return rc


_raw_stream = getattr(torch._C, "_cuda_getCurrentRawStream", None)


def stream_ptr(device=None):
    """Raw handle of torch's current stream on `device` (a torch.device, an
    index or None for the current device)."""
    if _raw_stream is not None:
        if device is None:
            idx = torch.cuda.current_device()
        elif isinstance(device, int):
            idx = device
        else:
            device = torch.device(device)
            idx = device.index if device.index is not None else torch.cuda.current_device()
        return _raw_stream(idx)
    return torch.cuda.current_stream(device).cuda_stream


def ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def dtype_code(dtype):
    if dtype == torch.float32:
        return TQ_F32
    if dtype == torch.float64:
        return TQ_F64
    raise ValueError(f"Unsupported dtype {dtype}; torchquad_amd computes in float32 or float64")


class _Workspace:
    """Per (device, stream) scratch buffers grown on demand, zeroed when
    allocated (the fused grid kernel keeps a completion counter in the first
    bytes that must start at zero; every launch leaves it zero).

    The kernels of one call finish reading a workspace before any later call
    on the same stream writes it, so reuse within a stream is safe.
    """

    def __init__(self):
        self._bufs = {}
        self._lock = threading.Lock()

    def get(self, nbytes, device, tag="main", stream=None):
        """The buffer of (device, stream, tag); stream: a raw stream handle
        already looked up by the caller (default: torch's current stream)."""
        if stream is None:
            stream = torch.cuda.current_stream(device).cuda_stream
        key = (device.index if device.index is not None else torch.cuda.current_device(), stream, tag)
        with self._lock:
            buf = self._bufs.get(key)
            if buf is None or buf.numel() < nbytes:
                size = max(int(nbytes), 256)
                size = 1 << (size - 1).bit_length()
                buf = torch.zeros(size, dtype=torch.uint8, device=device)
                self._bufs[key] = buf
            return buf


workspace = _Workspace()


def env_flag(name, default="0"):
    return os.environ.get(name, default) not in ("0", "", "false", "False")
