"""Global configuration helpers mirroring torchquad/utils (set_precision,
set_up_backend, enable_cuda, set_log_level)."""
from .set_precision import set_precision, _get_precision
from .set_up_backend import set_up_backend, _get_default_backend
from .enable_cuda import enable_cuda
from .set_log_level import set_log_level, logger

__all__ = [
    "set_precision",
    "set_up_backend",
    "enable_cuda",
    "set_log_level",
    "logger",
    "_get_precision",
    "_get_default_backend",
]
