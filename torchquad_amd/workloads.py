"""Seeded synthetic workloads of Genz-family integrals for the batched and
serving forms (numpy PCG64).

``genz_batch(count, seed)`` draws, per integral: dim ~ U{1..6}; total points
N log-uniform in [1e3, 1e6]; rule ~ U{trapezoid, simpson, gauss_legendre};
family ~ U{Gaussian peak, oscillatory, product peak}; a_k ~ U(0, 5),
u_k ~ U(0, 1), w ~ U(0, 1); domain [0, 1]^dim.
"""
import numpy as np

from .integrands import GaussianPeak, Oscillatory, ProductPeak

RULES = ("trapezoid", "simpson", "gauss_legendre")
FAMILIES = ("gaussian_peak", "oscillatory", "product_peak")


def genz_item(rng):
    dim = int(rng.integers(1, 7))
    N = int(round(10 ** rng.uniform(3.0, 6.0)))
    rule = RULES[int(rng.integers(0, 3))]
    fam = FAMILIES[int(rng.integers(0, 3))]
    a = rng.uniform(0.0, 5.0, dim)
    a = np.where(a < 1e-3, 1e-3, a)
    u = rng.uniform(0.0, 1.0, dim)
    w = float(rng.uniform(0.0, 1.0))
    if fam == "gaussian_peak":
        fn = GaussianPeak(a, u)
    elif fam == "oscillatory":
        fn = Oscillatory(a, w)
    else:
        fn = ProductPeak(a, u)
    return {"rule": rule, "fn": fn, "dim": dim, "N": N, "domain": [[0.0, 1.0]] * dim, "family": fam,
            "a": a, "u": u, "w": w}


def genz_batch(count, seed=0):
    rng = np.random.default_rng(seed)
    return [genz_item(rng) for _ in range(count)]
