"""Logging configuration (reference: torchquad/utils/set_log_level.py:5-18).

The reference configures loguru; this package logs through the standard
``logging`` module under the logger name ``torchquad_amd``.  Log calls on the
hot path never format device tensors, so logging never forces a device
synchronisation.
"""
import logging
import sys

logger = logging.getLogger("torchquad_amd")
if not logger.handlers:
    logger.addHandler(logging.NullHandler())

_LEVELS = {
    "TRACE": logging.DEBUG,
    "DEBUG": logging.DEBUG,
    "INFO": logging.INFO,
    "SUCCESS": logging.INFO,
    "WARNING": logging.WARNING,
    "ERROR": logging.ERROR,
    "CRITICAL": logging.CRITICAL,
}


def set_log_level(log_level: str):
    """Set the log level and attach a stderr handler (set_log_level.py:5-18)."""
    level = _LEVELS.get(str(log_level).upper())
    if level is None:
        raise ValueError(f"Unknown log level {log_level}")
    for h in list(logger.handlers):
        logger.removeHandler(h)
    handler = logging.StreamHandler(sys.stderr)
    handler.setFormatter(logging.Formatter("|TQ-%(levelname)s| %(message)s"))
    logger.addHandler(handler)
    logger.setLevel(level)
    logger.info(f"Setting LogLevel to {log_level}")
