"""Default floating point precision (reference: torchquad/utils/set_precision.py:6-89)."""
import os
import sys

from .set_log_level import logger


def _get_precision(backend):
    """Configured default precision from TORCHQUAD_DTYPE_<BACKEND> (set_precision.py:6-15)."""
    return os.environ.get(f"TORCHQUAD_DTYPE_{backend.upper()}", None)


def set_precision(data_type="float32", backend="torch"):
    """Set torch's default dtype (set_precision.py:18-89).

    "float"/"double" are accepted aliases; anything else falls back to
    float32 with an error message, as in the reference.  When the GPU has
    been initialised the default device becomes the GPU.
    """
    data_type = {"float": "float32", "double": "float64"}.get(data_type.lower(), data_type)
    if data_type not in ["float32", "float64"]:
        error_msg = (
            f'Invalid data type "{data_type}". Only float32 and float64 are supported. '
            "Setting the data type to float32."
        )
        logger.error(error_msg)
        print(f"ERROR: {error_msg}", file=sys.stderr)
        data_type = "float32"
    if backend == "torch":
        import torch

        torch.set_default_dtype({"float32": torch.float32, "float64": torch.float64}[data_type])
        if torch.cuda.is_initialized():
            torch.set_default_device("cuda")
            logger.info(f"Setting Torch's default dtype to {data_type} and device to CUDA.")
        else:
            logger.info(f"Setting Torch's default dtype to {data_type} (CPU).")
    elif backend == "numpy":
        os.environ["TORCHQUAD_DTYPE_NUMPY"] = data_type
        logger.info(f"NumPy default dtype set to {_get_precision('numpy')}")
    else:
        error_msg = f"Changing the data type is not supported for backend {backend}"
        logger.error(error_msg)
        print(f"ERROR: {error_msg}", file=sys.stderr)
