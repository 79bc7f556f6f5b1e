"""Backend selection (reference: torchquad/utils/set_up_backend.py:8-43).

Only the "torch" backend computes in this package (on the GPU through the
HIP library); the name is still recorded for compatibility.
"""
import os

from .enable_cuda import enable_cuda
from .set_precision import set_precision


def _get_default_backend():
    """Backend of the latest set_up_backend call, "torch" otherwise (set_up_backend.py:8-12)."""
    return os.environ.get("TORCHQUAD_DEFAULT_BACKEND", "torch")


def set_up_backend(backend, data_type=None, torch_enable_cuda=True):
    """Configure a numerical backend (set_up_backend.py:15-43)."""
    if backend == "torch":
        if torch_enable_cuda:
            if data_type is None:
                enable_cuda()
            else:
                enable_cuda(data_type=None)
    if data_type is not None:
        set_precision(data_type, backend=backend)
    os.environ["TORCHQUAD_DEFAULT_BACKEND"] = backend
