"""Pins the oracle (CPU): golden vectors, known answers and mathematics."""
import math

import numpy as np
import pytest
from scipy import integrate as sint

from oracle import families as fam
from oracle import grid, native


def test_philox_known_answers(ref_vectors):
    for case in ref_vectors["philox4x32_10_kat"]["cases"]:
        out = native.philox4x32_10(case["ctr"], case["key"])
        assert [int(v) for v in out] == case["out"]


def test_uniform_stream_definition():
    # word -> float mapping and block layout of the documented stream
    u32 = native.uniform(1234, 7, 3, 9, np.float32)
    assert u32.dtype == np.float32 and np.all(u32 >= 0) and np.all(u32 < 1)
    for r in range(3):
        for c in range(9):
            w = native.philox4x32_10([(7 + r) & 0xFFFFFFFF, 0, c // 4, 0], [1234, 0])
            assert u32[r, c] == np.float32((int(w[c % 4]) >> 8) * 2.0**-24)
    u64 = native.uniform(99, 0, 2, 3, np.float64)
    w = native.philox4x32_10([1, 0, 1, 0], [99, 0])
    expect = (((int(w[0]) >> 5) << 26) | (int(w[1]) >> 6)) * 2.0**-53
    assert u64[1, 2] == expect


def test_uniform_stream_statistics():
    u = native.uniform(5, 0, 20000, 4, np.float32).astype(np.float64)
    assert abs(u.mean() - 0.5) < 0.01
    assert abs(u.var() - 1.0 / 12.0) < 0.003
    c = np.corrcoef(u.T)
    assert np.max(np.abs(c - np.eye(4))) < 0.03


@pytest.mark.parametrize("case_i", range(5))
def test_grid_oracle_tutorial_outputs(ref_vectors, case_i):
    case = ref_vectors["tutorial_outputs"]["cases"][case_i]
    if case["integrand"] == "exp(x)*x^2":
        f = lambda x: np.exp(x[:, 0]) * x[:, 0] ** 2  # noqa: E731
    else:
        f = lambda x: np.sum(np.sin(x), axis=1)  # noqa: E731
    val = grid.integrate(case["rule"], f, case["dim"], case["N"], case["domain"])
    assert abs(val - case["value"]) <= ref_vectors["tutorial_outputs"]["rel_tol_fp32"] * abs(case["value"])


@pytest.mark.parametrize(
    "rule,N,max_deg",
    [("trapezoid", 2, 1), ("simpson", 3, 3), ("boole", 5, 5), ("gauss_legendre", 4, 7)],
)
def test_grid_oracle_polynomial_exactness(rule, N, max_deg):
    # trapezoid_test.py:34-48, simpson_test.py:36-50, boole_test.py:29-33, gauss_test.py:24-45
    rng = np.random.default_rng(1)
    for deg in range(max_deg + 1):
        coeffs = rng.uniform(-2, 2, deg + 1)
        a, b = -1.3, 2.1
        val = grid.integrate(rule, lambda x: np.polyval(coeffs[::-1], x[:, 0]), 1, N, [[a, b]])
        P = np.polynomial.polynomial.Polynomial(coeffs).integ()
        assert abs(val - (P(b) - P(a))) < 1e-12 * max(1.0, abs(P(b) - P(a)))
    # one degree more is not exact (guards against an over-accurate rule)
    coeffs = np.zeros(max_deg + 2)
    coeffs[-1] = 1.0
    val = grid.integrate(rule, lambda x: x[:, 0] ** (max_deg + 1), 1, N, [[0.0, 1.0]])
    assert abs(val - 1.0 / (max_deg + 2)) > 1e-6


@pytest.mark.parametrize("rule", ["trapezoid", "simpson", "boole", "gauss_legendre"])
def test_grid_oracle_three_formulations_agree(rule):
    rng = np.random.default_rng(2)
    d, n = 3, 9 if rule != "simpson" else 9
    dom = np.array([[-1.0, 0.5], [0.0, 2.0], [-0.5, 0.5]])
    a = rng.uniform(1, 3, d)
    u = rng.uniform(-0.5, 0.5, d)
    full = grid.integrate(rule, lambda x: fam.gaussian_peak(x, a, u), d, n**d, dom)
    sep = grid.separable_gaussian_peak(rule, a, u, dom, n)
    bf = native.grid_sum(rule, "gaussian_peak", d, n, dom, a=a, u=u)
    assert abs(full - sep) < 1e-14 * abs(sep)
    assert abs(bf - sep) < 1e-14 * abs(sep)
    c = rng.uniform(0, 2, d)
    w = 0.3
    full = grid.integrate(rule, lambda x: fam.oscillatory(x, c, w), d, n**d, dom)
    sep = grid.separable_oscillatory(rule, c, w, dom, n)
    bf = native.grid_sum(rule, "oscillatory", d, n, dom, c1=w, a=c)
    assert abs(full - sep) < 1e-13
    assert abs(bf - sep) < 1e-13


def test_brute_force_tiny_by_hand():
    # 2x2 trapezoid on [0,1]^2 of f=x0+2*x1 via exp_linear? use constant & linear by hand
    v = native.grid_sum("trapezoid", "constant", 2, 2, [[0, 1], [0, 3]], c0=2.5)
    assert v == pytest.approx(2.5 * 3.0, rel=1e-15)
    v = native.grid_sum("simpson", "exp_linear", 1, 3, [[0, 2]], a=[1.0])
    assert v == pytest.approx((1.0 + 4 * math.e + math.e**2) / 3.0, rel=1e-15)


def test_closed_forms_against_scipy():
    rng = np.random.default_rng(3)
    lo, hi = np.array([-0.7, 0.2]), np.array([0.4, 1.5])
    a, u = rng.uniform(1, 3, 2), rng.uniform(-0.5, 0.5, 2)

    def q(f):
        return sint.dblquad(lambda y, x: f(np.array([[x, y]]))[0], lo[0], hi[0], lo[1], hi[1],
                            epsabs=1e-13, epsrel=1e-12)[0]

    assert fam.gaussian_peak_integral(a, u, lo, hi) == pytest.approx(q(lambda x: fam.gaussian_peak(x, a, u)), rel=1e-10)
    c = rng.uniform(0, 2, 2)
    assert fam.oscillatory_integral(c, 0.37, lo, hi) == pytest.approx(q(lambda x: fam.oscillatory(x, c, 0.37)), rel=1e-9)
    assert fam.product_peak_integral(a, u, lo, hi) == pytest.approx(q(lambda x: fam.product_peak(x, a, u)), rel=1e-10)
    assert fam.exp_linear_integral(a, 0.2, lo, hi) == pytest.approx(q(lambda x: fam.exp_linear(x, a, 0.2)), rel=1e-10)
    assert fam.sin_sum_integral(a, u, lo, hi) == pytest.approx(q(lambda x: fam.sin_sum(x, a, u)), rel=1e-10)
    assert fam.double_gaussian_integral(2, lo, hi, 0.3) == pytest.approx(
        q(lambda x: fam.double_gaussian(x, 0.3)), rel=1e-10)


def test_configs_analytic(configs):
    assert configs["C1"]["analytic"] == pytest.approx(5.0732141117728515, rel=1e-15)
    c = configs["C2a"]
    sep = grid.separable_gaussian_peak("gauss_legendre", c["a"], c["u"], c["domain"], 64)
    assert sep == pytest.approx(c["analytic"], rel=1e-13)


@pytest.mark.parametrize("case_i", range(3))
def test_N_adjust(ref_vectors, case_i):
    case = ref_vectors["N_adjust"]["cases"][case_i]
    N = grid.adjust_N(case["rule"], case["dim"], case["N"])
    assert grid.n_per_dim(N, case["dim"]) == case["n"]


def test_mc_oracle_same_stream_mean():
    dom = [[0.0, 1.0]] * 3
    s1, s2 = native.mc_sums("float64", "constant", 3, 1000, 7, 0, dom)
    assert s1 == 1000.0 and s2 == 1000.0
    s1, _ = native.mc_sums("float64", "exp_linear", 3, 200000, 7, 0, dom, a=[1.0, 1.0, 1.0])
    assert s1 / 200000 == pytest.approx((math.e - 1) ** 3, rel=5e-3)


def test_torch_cpu_restatement_matches_oracle():
    """The torch-CPU restatement timed as bench.py's CPU baseline computes the
    same sums as the numpy oracle (FP64, small grids)."""
    import torch

    from oracle import families as fam
    from oracle import grid as og
    from oracle import torch_cpu

    a, u = np.array([2.0, 3.0, 1.5]), np.array([0.1, -0.2, 0.3])
    dom = [[-1.0, 1.0]] * 3
    for rule, n in (("boole", 9), ("simpson", 7), ("trapezoid", 6)):
        val, done, pts, _ = torch_cpu.grid_rule_slices(rule, torch_cpu.gaussian_peak_fn(a, u, torch.float64), dom, n,
                                                       dtype=torch.float64)
        ref = og.integrate(rule, lambda x: fam.gaussian_peak(x, a, u), 3, n**3, dom)
        assert done == n and pts == n**3
        assert abs(val - ref) <= 1e-13 * abs(ref)
    # a bounded run stops early and reports no value
    val, done, pts, _ = torch_cpu.grid_rule_slices("boole", torch_cpu.exp_sum_fn(), dom, 9, seconds=0.0)
    assert val is None and done == 1 and pts == 81


@pytest.mark.parametrize("family", ["gaussian_peak", "oscillatory", "product_peak"])
@pytest.mark.parametrize("rule,n", [("gauss_legendre", 7), ("trapezoid", 9), ("simpson", 9), ("boole", 9)])
def test_family_gradient_oracle(family, rule, n):
    """The torch-autograd restatement of the reference grid path: its value
    equals the separable oracle of the same rule, and its gradients equal
    central differences of that value (h = 1e-6, FP64)."""
    from oracle import grad as ogr

    rng = np.random.default_rng(zlib_seed(family + rule))
    d = 3
    p1 = rng.uniform(1.0, 3.0, d) * (rng.choice([-1.0, 1.0], d) if family == "oscillatory" else 1.0)
    p2 = rng.uniform(-0.3, 0.3, d) if family != "oscillatory" else float(rng.uniform(0, 1))
    lo = rng.uniform(-1.0, -0.2, d)
    hi = lo + rng.uniform(0.8, 1.5, d)
    dom = np.stack([lo, hi], 1)
    I, g1, g2, glo, ghi = ogr.reference_autograd_family(family, p1, p2, lo, hi, n, rule)
    if family == "gaussian_peak":
        ref = grid.separable_gaussian_peak(rule, p1, p2, dom, n)
    elif family == "oscillatory":
        ref = grid.separable_oscillatory(rule, p1, p2, dom, n)
    else:
        ref = grid.separable_product(rule, [lambda x, a=a, u=u: 1 / (a ** -2 + (x - u) ** 2) for a, u in zip(p1, p2)],
                                     dom, n)
    assert abs(I - ref) <= 1e-14 * max(abs(ref), 1.0)

    def value(q1, q2, l, h):
        return ogr.reference_autograd_family(family, q1, q2, l, h, n, rule)[0]

    eps = 1e-6
    for name, base, got in (("p1", p1, g1), ("p2", p2, g2), ("lo", lo, glo), ("hi", hi, ghi)):
        base = np.atleast_1d(np.asarray(base, dtype=np.float64))
        fd = np.zeros_like(base)
        for i in range(base.size):
            args = {"p1": p1, "p2": p2, "lo": lo, "hi": hi}
            plus, minus = base.copy(), base.copy()
            plus[i] += eps
            minus[i] -= eps
            args_p, args_m = dict(args), dict(args)
            args_p[name] = plus if np.ndim(args[name]) else float(plus[0])
            args_m[name] = minus if np.ndim(args[name]) else float(minus[0])
            fd[i] = (value(args_p["p1"], args_p["p2"], args_p["lo"], args_p["hi"])
                     - value(args_m["p1"], args_m["p2"], args_m["lo"], args_m["hi"])) / (2 * eps)
        np.testing.assert_allclose(np.atleast_1d(got), fd, rtol=1e-6, atol=1e-8, err_msg=f"{family} {rule} {name}")


def zlib_seed(name):
    import zlib

    return zlib.crc32(name.encode())
