"""Select the ROCm GPU as torch's default device (reference: torchquad/utils/enable_cuda.py:7-26)."""
import warnings

from .set_log_level import logger
from .set_precision import set_precision


def enable_cuda(data_type="float32"):
    """Make the GPU torch's default device and optionally set the precision."""
    import torch

    if torch.cuda.is_available():
        torch.cuda.init()
        logger.info("PyTorch VERSION: " + str(torch.__version__))
        logger.info("Number of ROCm devices: " + str(torch.cuda.device_count()))
        logger.info("Active device: GPU" + str(torch.cuda.current_device()))
        if data_type is not None:
            set_precision(data_type)
        else:
            torch.set_default_device("cuda")
    else:
        warning_msg = "Error enabling CUDA. cuda.is_available() returned False. CPU will be used."
        logger.warning(warning_msg)
        warnings.warn(warning_msg, RuntimeWarning)
