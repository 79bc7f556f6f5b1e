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


RULE_CODES = {"trapezoid": 0, "simpson": 1, "boole": 2, "gauss_legendre": 3}
FAMILY_CODES = {"gaussian_peak": 0, "oscillatory": 1, "product_peak": 2}


def genz_arrays(count, seed=0):
    """The integrals of genz_batch(count, seed) as arrays in the descriptor
    convention of tq_integrand / tq_batch_item (the form
    IntegrationServer.submit_many takes): rule, family, dim, N [count];
    domain [count, 6, 2]; a, u [count, 16]; c [count, 4] = (c0, c1, c2, c3)."""
    items = genz_batch(count, seed)
    out = {
        "rule": np.array([RULE_CODES[it["rule"]] for it in items], dtype=np.int64),
        "family": np.array([FAMILY_CODES[it["family"]] for it in items], dtype=np.int64),
        "dim": np.array([it["dim"] for it in items], dtype=np.int64),
        "N": np.array([it["N"] for it in items], dtype=np.int64),
        "domain": np.zeros((count, 6, 2)),
        "a": np.zeros((count, 16)),
        "u": np.zeros((count, 16)),
        "c": np.zeros((count, 4)),
    }
    for i, it in enumerate(items):
        d = it["dim"]
        out["domain"][i, :d, 1] = 1.0
        out["c"][i, 0] = 1.0
        if it["family"] == "oscillatory":
            out["a"][i, :d] = it["a"]
            out["c"][i, 1] = it["w"]
        else:
            out["a"][i, :d] = it["a"]
            out["u"][i, :d] = it["u"]
    return out
