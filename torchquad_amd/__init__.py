"""torchquad_amd: MI355X-native integration hot path with torchquad's API.

The integrators below keep the reference's names, signatures, defaults,
N-rounding, warnings and errors (esa/torchquad 0.5.0, torchquad/__init__.py
:10-48); their compute runs in hand-written HIP kernels for gfx950 loaded
from ``libtorchquad_amd.so`` through a C ABI (include/torchquad_amd.h).
"""
import os

__version__ = "0.5.0+mi355x"

TORCHQUAD_DISABLE_LOGGING = True

from .integration.integration_grid import IntegrationGrid
from .integration.monte_carlo import MonteCarlo
from .integration.trapezoid import Trapezoid
from .integration.simpson import Simpson
from .integration.boole import Boole
from .integration.gaussian import GaussHermite, GaussLegendre, Gaussian
from .integration.vegas import VEGAS
from .integration.vegas_map import VEGASMap
from .integration.vegas_stratification import VEGASStratification
from .integration.grid_integrator import GridIntegrator
from .integration.base_integrator import BaseIntegrator
from .integration.rng import RNG
from .utils.set_log_level import set_log_level
from .utils.enable_cuda import enable_cuda
from .utils.set_precision import set_precision
from .utils.set_up_backend import set_up_backend
from . import integrands
from .integrands import (
    GaussianPeak,
    Oscillatory,
    ProductPeak,
    ExpLinear,
    SinSum,
    DoubleGaussian,
    Constant,
)

__all__ = [
    "__version__",
    "GridIntegrator",
    "BaseIntegrator",
    "IntegrationGrid",
    "MonteCarlo",
    "Trapezoid",
    "Simpson",
    "Boole",
    "VEGAS",
    "VEGASMap",
    "VEGASStratification",
    "GaussLegendre",
    "GaussHermite",
    "Gaussian",
    "RNG",
    "enable_cuda",
    "set_precision",
    "set_log_level",
    "set_up_backend",
    "integrands",
    "GaussianPeak",
    "Oscillatory",
    "ProductPeak",
    "ExpLinear",
    "SinSum",
    "DoubleGaussian",
    "Constant",
]

if not TORCHQUAD_DISABLE_LOGGING:
    set_log_level(os.environ.get("TORCHQUAD_LOG_LEVEL", "WARNING"))
