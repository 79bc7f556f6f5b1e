"""Newton-Cotes base class (reference: torchquad/integration/newton_cotes.py:4-8)."""
from .grid_integrator import GridIntegrator


class NewtonCotes(GridIntegrator):
    """The abstract integrator that composite Newton-Cotes integrators inherit from."""

    def __init__(self):
        super().__init__()
