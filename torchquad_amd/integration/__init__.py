"""Integrators mirroring torchquad/integration on the HIP kernels."""
