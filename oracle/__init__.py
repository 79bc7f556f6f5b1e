"""ORACLE — test infrastructure only.

CPU restatements of the reference (esa/torchquad 0.5.0) algorithms on the
integration hot path, used as the checker by tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg.  Nothing in the torchquad_amd package
imports, links or calls anything here.

Modules:
  families  FP64 numpy definitions of the built-in integrand families and
            their closed-form integrals over boxes.
  grid      Newton-Cotes / Gauss-Legendre restatement (materialised grid,
            per-axis composite collapse) and separable FP64 product sums.
  gl        Gauss-Legendre node/weight pairs by mpmath root refinement.
  grad      erf closed-form integral and derivatives of the Gaussian peak
            over a box; torch-CPU FP64 autograd restatement of the
            reference's Gauss-Legendre path.
  vegas     VEGAS map / stratification / driver restatement.
  native    ctypes binding of liboracle.so (oracle.c: FP64 brute force
            tensor-product sums, Philox-4x32-10 and the uniform stream
            built on it, Monte Carlo sums mc_sums of
            monte_carlo.py:60-106).

Pinning status of each function is listed in the module docstrings; every
function is pinned by at least one check in tests/test_oracle.py unless its
docstring says "parity unpinned".
"""
