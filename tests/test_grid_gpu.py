"""Grid rules on the GPU vs the oracle (FP64 C brute force / numpy restatement).

Bars: |value - oracle| <= rtol * sum_p |w_p f(x_p)| with rtol 1e-6 (FP32) or
1e-12 (FP64).  For the positive families the sum of |w f| is |oracle|, so the
bar is a plain relative one; for sign-changing families (oscillatory, sin
sum) it is the scale of the rounding error of any summation order.  The
oracle evaluates on the inputs the kernel sees: an FP32 domain is rounded to
FP32 before the FP64 oracle runs on it.
"""
import math
import warnings
import zlib

import numpy as np
import pytest
import torch

import torchquad_amd as tq
from oracle import families as fam
from oracle import grid as og
from oracle import native

pytestmark = pytest.mark.gpu

RULES = {"trapezoid": tq.Trapezoid, "simpson": tq.Simpson, "boole": tq.Boole, "gauss_legendre": tq.GaussLegendre}
RTOL = {torch.float32: 1e-6, torch.float64: 1e-12}


@pytest.fixture(autouse=True)
def _quiet():
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        yield


def make_family(name, d, rng):
    a = rng.uniform(1.0, 5.0, d)
    u = rng.uniform(-0.5, 0.5, d)
    c = rng.uniform(0.0, 2.0, d)
    w = float(rng.uniform())
    if name == "gaussian_peak":
        return tq.GaussianPeak(a, u), dict(a=a, u=u), 1.0
    if name == "oscillatory":
        return tq.Oscillatory(c, w), dict(c1=w, a=c), 1.0
    if name == "product_peak":
        a = rng.uniform(0.5, 2.0, d)
        return tq.ProductPeak(a, u), dict(a=a, u=u), 1.0
    if name == "exp_linear":
        return tq.ExpLinear(d, a=c, c=0.1), dict(a=c, c1=0.1), 1.0
    if name == "sin_sum":
        return tq.SinSum(d, a=c, u=u), dict(a=c, u=u), float(d)
    if name == "double_gaussian":
        f = tq.DoubleGaussian(d, width=0.3)
        return f, dict(c0=f.norm, c1=1 / 3, c2=2 / 3, c3=1 / 0.09), f.norm * 2
    if name == "constant":
        return tq.Constant(d, 2.5), dict(c0=2.5), 2.5
    raise ValueError(name)


def domain_for(name, d, rng):
    if name == "double_gaussian":
        return [[0.0, 1.0]] * d
    lo = rng.uniform(-1.0, 0.0, d)
    return [[float(l), float(l + rng.uniform(0.5, 1.5))] for l in lo]


POSITIVE = {"gaussian_peak", "product_peak", "exp_linear", "double_gaussian", "constant"}


def seed_of(*key):
    """Fixed seed per case (zlib.crc32 does not depend on PYTHONHASHSEED)."""
    return zlib.crc32("/".join(map(str, key)).encode())


def as_seen(dom, dtype):
    """The domain as the kernel receives it (rounded to the tensor dtype)."""
    np_t = np.float32 if dtype == torch.float32 else np.float64
    return np.asarray(dom, dtype=np_t).astype(np.float64).tolist()


def oracle_value(rule, fname, d, n, dom, kw):
    """(sum w f, sum |w f|) by FP64 brute force."""
    return native.grid_sum(rule, fname, d, n, dom, with_abs=True, **kw)


def check(val, ref, dtype, scale, parity=None):
    """|val - ref| <= rtol * scale; scale is sum |w f| (= |ref| for positive integrands)."""
    tol = RTOL[dtype] * scale
    err = abs(val - ref)
    if parity is not None:
        parity.record(err / scale if scale else err, RTOL[dtype])
    assert err <= tol, f"got {val!r} ref {ref!r} diff {err:.3e} tol {tol:.3e}"


FAMS = ["gaussian_peak", "oscillatory", "product_peak", "exp_linear", "sin_sum", "double_gaussian", "constant"]
SHAPES = [(1, 33), (2, 17), (3, 13), (4, 9), (5, 9), (6, 5)]


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("rule", list(RULES))
@pytest.mark.parametrize("fname", FAMS)
def test_fused_vs_oracle(dtype, rule, fname, dev, parity):
    rng = np.random.default_rng(seed_of(rule, fname))
    for d, n in SHAPES:
        if rule == "boole":
            n = 4 * ((n - 1) // 4) + 1 if n >= 5 else 5
        if rule == "simpson" and n % 2 == 0:
            n += 1
        f, kw, _ = make_family(fname, d, rng)
        dom = domain_for(fname, d, rng)
        ref, scale = oracle_value(rule, fname, d, n, as_seen(dom, dtype), kw)
        if fname in POSITIVE:
            assert scale == pytest.approx(abs(ref), rel=1e-12)
        domain = torch.tensor(dom, dtype=dtype, device=dev)
        val = RULES[rule]().integrate(f, d, N=n**d, integration_domain=domain)
        assert val.dtype == dtype and val.shape == ()
        check(float(val), ref, dtype, scale, parity)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("rule", list(RULES))
def test_user_callable_vs_oracle(dtype, rule, dev, parity):
    rng = np.random.default_rng(11)
    for d, n in [(1, 21), (2, 9), (3, 9), (4, 5)]:
        if rule == "simpson" and n % 2 == 0:
            n += 1
        f, kw, _ = make_family("gaussian_peak", d, rng)
        dom = domain_for("gaussian_peak", d, rng)
        ref, scale = oracle_value(rule, "gaussian_peak", d, n, as_seen(dom, dtype), kw)
        domain = torch.tensor(dom, dtype=dtype, device=dev)
        integrator = RULES[rule]()
        val = integrator.integrate(lambda x: f(x), d, N=n**d, integration_domain=domain)
        assert integrator._nr_of_fevals == n**d
        check(float(val), ref, dtype, scale, parity)
        fused = RULES[rule]().integrate(f, d, N=n**d, integration_domain=domain)
        check(float(fused), ref, dtype, scale, parity)


def test_baseline_boole_config(configs, dev):
    c = configs["C2a"]
    f = tq.GaussianPeak(c["a"], c["u"])
    dom = torch.tensor(c["domain"], dtype=torch.float32, device=dev)
    for n in (17, 33, 65):
        val = float(tq.Boole().integrate(f, 5, N=n**5, integration_domain=dom))
        ref = og.separable_gaussian_peak("boole", c["a"], c["u"], c["domain"], n)
        assert abs(val - ref) <= 1e-6 * abs(ref), (n, val, ref)
    # n=65 Boole is within 1e-6 of the analytic value
    assert abs(val - c["analytic"]) <= 1e-6 * c["analytic"]


def test_baseline_c1_config(configs, dev):
    c = configs["C1"]
    dom = torch.tensor(c["domain"], dtype=torch.float64, device=dev)
    f = tq.ExpLinear(3)
    for rule in ("trapezoid", "simpson"):
        val = float(RULES[rule]().integrate(f, 3, N=c["N"], integration_domain=dom))
        ref = og.separable_product(rule, [np.exp] * 3, c["domain"], 101)
        assert abs(val - ref) <= 1e-12 * ref
    assert abs(val - c["analytic"]) <= 2e-10 * c["analytic"]


@pytest.mark.parametrize("case_i", range(5))
def test_tutorial_outputs(ref_vectors, case_i, dev):
    case = ref_vectors["tutorial_outputs"]["cases"][case_i]
    if case["integrand"] == "exp(x)*x^2":
        f = lambda x: torch.exp(x[:, 0]) * x[:, 0] ** 2  # noqa: E731
    else:
        f = lambda x: torch.sum(torch.sin(x), dim=1)  # noqa: E731
    dom = torch.tensor(case["domain"], dtype=torch.float32, device=dev)
    val = float(RULES[case["rule"]]().integrate(f, case["dim"], N=case["N"], integration_domain=dom))
    assert abs(val - case["value"]) <= 1e-6 * abs(case["value"])
    if case["integrand"] == "sum sin(x)":
        fused = float(RULES[case["rule"]]().integrate(tq.SinSum(case["dim"]), case["dim"], N=case["N"],
                                                    integration_domain=dom))
        assert abs(fused - case["value"]) <= 1e-6 * abs(case["value"])


@pytest.mark.parametrize("rule,N,max_deg", [("trapezoid", 2, 1), ("simpson", 3, 3), ("boole", 5, 5),
                                            ("gauss_legendre", 4, 7)])
def test_polynomial_exactness(rule, N, max_deg, dev):
    rng = np.random.default_rng(4)
    dom = torch.tensor([[-1.3, 2.1]], dtype=torch.float64, device=dev)
    for deg in range(max_deg + 1):
        coeffs = rng.uniform(-2, 2, deg + 1)
        ct = torch.tensor(coeffs, device=dev)
        val = float(RULES[rule]().integrate(lambda x: sum(ct[k] * x[:, 0] ** k for k in range(deg + 1)), 1, N=N,
                                            integration_domain=dom))
        P = np.polynomial.polynomial.Polynomial(coeffs).integ()
        exact = P(2.1) - P(-1.3)
        assert abs(val - exact) < 1e-13 * max(1.0, abs(exact))


def test_deterministic(configs, dev):
    c = configs["C2a"]
    f = tq.GaussianPeak(c["a"], c["u"])
    dom = torch.tensor(c["domain"], dtype=torch.float32, device=dev)
    vals = [tq.Boole().integrate(f, 5, N=33**5, integration_domain=dom).item() for _ in range(3)]
    assert vals[0] == vals[1] == vals[2]


def test_split_api_matches_integrate(dev):
    # trapezoid_test.py:141-174
    dom = torch.tensor([[0.0, 1.0]], device=dev)
    integ = tq.Trapezoid()
    pts, hs, n = integ.calculate_grid(101, dom)
    assert pts.shape == (101, 1) and n == 101
    vals, _ = integ.evaluate_integrand(lambda x: torch.exp(x[:, 0]), pts)
    r1 = integ.calculate_result(vals, 1, n, hs, dom)
    r2 = integ.calculate_result(function_values=vals, dim=1, n_per_dim=n, hs=hs, integration_domain=dom)
    r3 = integ.calculate_result(vals, dim=1, n_per_dim=n, hs=hs, integration_domain=dom)
    assert torch.allclose(r1, r2, rtol=1e-10) and torch.allclose(r1, r3, rtol=1e-10)
    assert abs(float(r1) - og.integrate("trapezoid", lambda x: np.exp(x[:, 0]), 1, 101, [[0, 1]])) < 1e-6
    # Gauss-Legendre split API: values are pre-weighted by evaluate_integrand
    gl = tq.GaussLegendre()
    dom64 = torch.tensor([[0.0, 2.0], [-1.0, 1.0]], dtype=torch.float64, device=dev)
    pts, hs, n = gl.calculate_grid(8**2, dom64)
    vals, _ = gl.evaluate_integrand(lambda x: torch.exp(x[:, 0]) * x[:, 1] ** 2, pts, weights=gl._weights(n, 2, "torch"))
    r = gl.calculate_result(vals, 2, n, hs, dom64)
    assert abs(float(r) - (math.e**2 - 1) * 2 / 3) < 1e-12


def test_grid_points_layout(dev):
    dom = torch.tensor([[0.0, 2.0], [-2.0, 1.0], [0.5, 1.0]], dtype=torch.float64, device=dev)
    g = tq.IntegrationGrid(4**3, dom)
    ref, axes = og.grid_points("trapezoid", dom.cpu().numpy(), 4)
    assert g._N == 4 and g.points.shape == (64, 3)
    np.testing.assert_allclose(g.points.cpu().numpy(), ref, rtol=0, atol=1e-15)
    np.testing.assert_allclose(g.h.cpu().numpy(), [ax[1] - ax[0] for ax in axes], rtol=1e-15)


def test_multidim_integrand(dev):
    dom = torch.tensor([[0.0, 1.0], [0.0, 2.0]], dtype=torch.float64, device=dev)

    def f(x):
        base = torch.exp(x[:, 0] + 0.5 * x[:, 1])
        return torch.stack([torch.stack([base, 2 * base], 1), torch.stack([x[:, 0], x[:, 1] ** 2], 1)], 1)

    val = tq.Simpson().integrate(f, 2, N=21**2, integration_domain=dom)
    assert val.shape == (2, 2)
    ref = og.integrate("simpson", lambda x: np.stack([np.stack([np.exp(x[:, 0] + 0.5 * x[:, 1]),
                                                                  2 * np.exp(x[:, 0] + 0.5 * x[:, 1])], 1),
                                                        np.stack([x[:, 0], x[:, 1] ** 2], 1)], 1),
                       2, 21**2, [[0, 1], [0, 2]])
    np.testing.assert_allclose(val.cpu().numpy(), ref, rtol=1e-12)


def test_gradients_reference_answers(ref_vectors, dev):
    G = ref_vectors["gradients"]
    for integ, N1, N2 in zip([tq.Trapezoid(), tq.Simpson(), tq.Boole()], G["Ns_1d"], G["Ns_2d"]):
        dom = torch.tensor(G["v_function_domain"]["domain"], dtype=torch.float64, device=dev, requires_grad=True)
        val = integ.integrate(lambda x: 2 * torch.abs(x), 1, N=N1, integration_domain=dom)
        val.backward()
        assert abs(float(val) - 2.0) < 1e-2
        np.testing.assert_array_less(np.abs(dom.grad.cpu().numpy() - np.array([[-2.0, 2.0]])), 2e-2)
        dom = torch.tensor(G["poly_domain"]["domain"], dtype=torch.float64, device=dev, requires_grad=True)
        val = integ.integrate(lambda x: 1.0 + 2.0 * x[:, 0] + 3.0 * x[:, 1] ** 2, 2, N=N2, integration_domain=dom)
        val.backward()
        assert abs(float(val) - 12.0) < 8e-2
        np.testing.assert_array_less(np.abs(dom.grad.cpu().numpy() - np.array(G["poly_domain"]["grad"])), 0.1)
        param = torch.tensor([1.0, 2.0, 3.0], dtype=torch.float64, device=dev, requires_grad=True)
        val = integ.integrate(lambda x: param[0] + param[1] * x[:, 0] + param[2] * x[:, 1] ** 2, 2, N=N2,
                              integration_domain=[[0.0, 1.0], [0.0, 2.0]])
        val.backward()
        np.testing.assert_array_less(np.abs(param.grad.cpu().numpy() - np.array([2.0, 1.0, 8.0 / 3.0])), 5e-2)
    # Gauss-Legendre (never gradient-tested by the reference: gradient_test.py:168-178)
    gl = tq.GaussLegendre()
    dom = torch.tensor([[0.0, 1.0], [0.0, 2.0]], dtype=torch.float64, device=dev, requires_grad=True)
    val = gl.integrate(lambda x: 1.0 + 2.0 * x[:, 0] + 3.0 * x[:, 1] ** 2, 2, N=8**2, integration_domain=dom)
    val.backward()
    assert abs(float(val) - 12.0) < 1e-12
    np.testing.assert_allclose(dom.grad.cpu().numpy(), [[-10.0, 14.0], [-2.0, 14.0]], atol=1e-11)


def test_gl_generator_matches_leggauss(dev):
    from torchquad_amd.gl_nodes import gauss_legendre

    from oracle import gl as ogl

    for n in (1, 2, 5, 32, 255, 600):
        x, w = gauss_legendre(n, dev)
        rx, rw = np.polynomial.legendre.leggauss(n)
        np.testing.assert_allclose(x.cpu().numpy(), rx, atol=2e-15 * max(1, n / 50))
        # leggauss itself carries ~1e-10 relative error in the endpoint weights
        np.testing.assert_allclose(w.cpu().numpy(), rw, rtol=2e-10)
        xs, ws = x.cpu().numpy(), w.cpu().numpy()
        for k in sorted({0, min(1, n - 1), n // 3, n - 1}):
            mx, mw = ogl.node_weight(n, k)
            assert abs(xs[k] - mx) <= 1e-15, (n, k)
            assert abs(ws[k] - mw) <= 2e-12 * mw, (n, k)
    x, w = gauss_legendre(100000, dev)
    assert abs(float(w.sum()) - 2.0) < 1e-11
    assert float((x[1:] - x[:-1]).min()) > 0
    assert abs(float(torch.sum(w * x**2)) - 2.0 / 3.0) < 1e-11


def test_large_one_dimensional(dev):
    dom = torch.tensor([[0.0, 3.0]], dtype=torch.float64, device=dev)
    f = tq.ExpLinear(1, a=[-0.7])
    for rule, n in (("trapezoid", 1_000_000), ("simpson", 999_999), ("gauss_legendre", 20_000)):
        val = float(RULES[rule]().integrate(f, 1, N=n, integration_domain=dom))
        ref = og.separable_product(rule, [lambda x: np.exp(-0.7 * x)], [[0.0, 3.0]], n) if rule != "gauss_legendre" \
            else fam.exp_linear_integral([-0.7], 0.0, [0.0], [3.0])
        assert abs(val - ref) <= 1e-12 * abs(ref), rule


def test_max_dim(dev):
    d = 16
    dom = torch.tensor([[0.0, 1.0]] * d, dtype=torch.float64, device=dev)
    val = float(tq.Trapezoid().integrate(tq.ExpLinear(d, a=[0.1] * d), d, N=2**d, integration_domain=dom))
    ref = og.separable_product("trapezoid", [lambda x: np.exp(0.1 * x)] * d, [[0, 1]] * d, 2)
    assert abs(val - ref) <= 1e-12 * ref


@pytest.mark.parametrize("cls", [tq.Boole, tq.GaussLegendre])
def test_compiled_integrate_graph_replays(cls, dev):
    """get_jit_compiled_integrate: graph replays with new domains equal the eager results."""
    f = tq.GaussianPeak([2.0, 3.0, 1.5], [0.1, -0.2, 0.3])
    integ = cls()
    comp = integ.get_jit_compiled_integrate(3, N=21**3, integration_domain=[[-1.0, 1.0]] * 3)
    rng = np.random.default_rng(3)
    for rep in range(6):
        lo = rng.uniform(-1.0, 0.0, 3)
        dom = torch.tensor(np.stack([lo, lo + rng.uniform(0.5, 2.0, 3)], 1), dtype=torch.float32, device=dev)
        got = comp(f, dom)
        # eager work and a host synchronisation between replays
        ref = cls().integrate(f, 3, N=21**3, integration_domain=dom)
        junk = torch.randn(4096, device=dev).sum()
        torch.cuda.synchronize()
        assert got.shape == ref.shape and float(got) == float(ref), (rep, float(got), float(ref))
        assert float(junk) == float(junk)
    # another built-in integrand gets its own graph; a user callable runs eagerly
    g = tq.Oscillatory([1.0, 2.0, 0.5], 0.25)
    dom = torch.tensor([[0.0, 1.0]] * 3, dtype=torch.float32, device=dev)
    assert float(comp(g, dom)) == float(cls().integrate(g, 3, N=21**3, integration_domain=dom))
    h = lambda x: torch.sum(x * x, dim=1)  # noqa: E731
    assert float(comp(h, dom)) == pytest.approx(1.0, rel=1e-5)


def test_gauss_hermite_rule(dev):
    """Gauss-Hermite on the fused and point/reduce kernels vs the tutorial's value, closed forms
    and the reference-style subclass (tutorial.rst:1198-1236) on the reference layout."""
    gh = tq.GaussHermite()
    # the tutorial's example: int exp(-x^2) (1 - x) dx = sqrt(pi)
    val = float(gh.integrate(lambda x: 1 - x, dim=1, N=200,
                             integration_domain=torch.tensor([[-1.0, 1.0]], dtype=torch.float64, device=dev)))
    assert abs(val - math.sqrt(math.pi)) <= 1e-12 * math.sqrt(math.pi)
    dom2 = torch.tensor([[-1.0, 1.0]] * 2, dtype=torch.float64, device=dev)
    # fused: int exp(-|x|^2) dx over R^2 = pi; int exp(-x^2 + x/2) dx = sqrt(pi) e^(1/16)
    assert abs(float(gh.integrate(tq.Constant(2, 1.0), 2, N=30**2, integration_domain=dom2)) - math.pi) <= 1e-12 * math.pi
    dom1 = torch.tensor([[-1.0, 1.0]], dtype=torch.float64, device=dev)
    ref = math.sqrt(math.pi) * math.exp(1.0 / 16.0)
    assert abs(float(gh.integrate(tq.ExpLinear(1, [0.5]), 1, N=40, integration_domain=dom1)) - ref) <= 1e-12 * ref

    class UserHermite(tq.Gaussian):  # the tutorial's subclass: classified, then run on the kernels
        def __init__(self):
            super().__init__()
            import scipy.special

            self._root_fn = scipy.special.roots_hermite

        @staticmethod
        def _apply_composite_rule(cur_dim_areas, dim, hs, domain):
            for _ in range(dim):
                cur_dim_areas = torch.sum(cur_dim_areas, dim=len(cur_dim_areas.shape) - 1)
            return cur_dim_areas

    from torchquad_amd.integration.gaussian import UNIT

    uh = UserHermite()
    assert uh._tq_mode() == UNIT
    f = tq.GaussianPeak([0.7, 1.3], [0.2, -0.1])
    a = float(gh.integrate(f, 2, N=24**2, integration_domain=dom2))
    b = float(uh.integrate(f, 2, N=24**2, integration_domain=dom2))
    # same tables, same kernels: identical
    assert a == b
    # against the rule's own FP64 sum on the host (the reference layout's arithmetic)
    import scipy.special

    r, w = scipy.special.roots_hermite(24)
    X, Y = np.meshgrid(r, r, indexing="ij")
    ref = float(np.sum(np.outer(w, w) * np.exp(-(0.7 * (X - 0.2)) ** 2 - (1.3 * (Y + 0.1)) ** 2)))
    assert abs(b - ref) <= 1e-12 * abs(ref)
    # user callables on the point/reduce kernels, the domain's values ignored as in the reference
    c = float(uh.integrate(lambda x: f(x), 2, N=24**2, integration_domain=dom2 * 3.0))
    assert abs(c - ref) <= 1e-12 * abs(ref)
    # the split API: points are the unresized roots, the result a plain sum
    pts, hs, n = uh.calculate_grid(24**2, dom2 * 5.0)
    assert n == 24 and abs(float(pts[0, 0]) - r[0]) <= 1e-15
    vals, _ = uh.evaluate_integrand(lambda x: f(x), pts, weights=uh._weights(n, 2, "torch"))
    assert abs(float(uh.calculate_result(vals, 2, n, hs, dom2 * 5.0)) - ref) <= 1e-12 * abs(ref)


def test_custom_root_fn_rule_fused(dev, parity):
    """A rule that only swaps _root_fn (gaussian.py:13-25): its roots and
    weights run on the fused kernels with Gauss-Legendre's map and scale.
    Here the nodes come from scipy.special.roots_legendre instead of numpy's
    leggauss; against the FP64 oracle, and through the compiled graph."""
    import scipy.special

    class ScipyLegendre(tq.GaussLegendre):
        def __init__(self):
            super().__init__()
            self._root_fn = scipy.special.roots_legendre

    from torchquad_amd.integration.gaussian import AFFINE

    g = ScipyLegendre()
    assert g._tq_mode() == AFFINE
    f = tq.GaussianPeak([2.0, 3.0, 1.5], [0.1, -0.2, 0.3])
    dom = [[-1.0, 0.5], [0.0, 1.0], [-0.5, 0.7]]
    for dtype in (torch.float32, torch.float64):
        val = float(g.integrate(f, 3, N=20**3, integration_domain=torch.tensor(dom, dtype=dtype, device=dev)))
        ref, scale = oracle_value("gauss_legendre", "gaussian_peak", 3, 20, as_seen(dom, dtype), dict(a=f.a, u=f.u))
        check(val, ref, dtype, scale, parity)
    # graph-replayed compiled integrate on the same custom rule
    comp = g.get_jit_compiled_integrate(3, N=20**3, integration_domain=dom)
    d32 = torch.tensor(dom, dtype=torch.float32, device=dev)
    assert float(comp(f, d32)) == float(g.integrate(f, 3, N=20**3, integration_domain=d32))


SPLIT_N = {
    "trapezoid": (128, 129, 191, 192, 193, 255, 256, 257, 1000, 4097),
    "simpson": (129, 191, 193, 255, 257, 1001, 4097),
    "boole": (129, 193, 257, 1001, 4097),
    "gauss_legendre": (128, 129, 192, 255, 257, 1000, 4097),
}


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("rule", list(RULES))
def test_split_mode_sizes_vs_oracle(dtype, rule, dev, parity):
    """One axis of n >= 128 nodes is walked in rows of 64 nodes: Newton-Cotes interior rows in a tight
    loop with one weight per lane, the rows holding the end nodes (and nodes past n) in the general
    body; Gauss-Legendre rows over the lower half of the table, each node with its mirror (odd n:
    the middle node once).  Sizes around the row boundaries, every family, against the FP64 brute
    force."""
    rng = np.random.default_rng(17)
    for n in SPLIT_N[rule]:
        for fname in FAMS:
            f, kw, _ = make_family(fname, 1, rng)
            dom = domain_for(fname, 1, rng)
            ref, scale = oracle_value(rule, fname, 1, n, as_seen(dom, dtype), kw)
            val = RULES[rule]().integrate(f, 1, N=n, integration_domain=torch.tensor(dom, dtype=dtype, device=dev))
            check(float(val), ref, dtype, scale, parity)


MERGE_SHAPES = [(3, 61), (4, 40), (4, 65), (5, 33)]


def separable_ref(rule, fname, f, dom, n):
    """FP64 sum of the rule for the product-form families (and sums of products)."""
    d = len(dom)
    if fname == "gaussian_peak":
        return og.separable_gaussian_peak(rule, f.a, f.u, dom, n)
    if fname == "oscillatory":
        return og.separable_oscillatory(rule, f.c, f.w, dom, n)
    if fname == "product_peak":
        return og.separable_product(rule, [lambda x, a=a, u=u: 1.0 / (a**-2.0 + (x - u) ** 2)
                                           for a, u in zip(f.a, f.u)], dom, n)
    inv = 1.0 / (f.width * f.width)
    return f.norm * sum(og.separable_product(rule, [lambda x, c=c: np.exp(-inv * (x - c) ** 2)] * d, dom, n)
                        for c in f.centers)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("rule", list(RULES))
def test_two_axis_loops_vs_oracle(dtype, rule, dev, parity):
    """Shapes whose loop runs over two axes (n^2-entry tables; above 16 KB of tables the blocks are
    8 waves), for one- and two-slot families, against separable FP64 sums of the same rule."""
    rng = np.random.default_rng(23)
    for d, n in MERGE_SHAPES:
        if rule in ("simpson", "boole") and n % 2 == 0:
            n += 1
        for fname in ("gaussian_peak", "oscillatory", "product_peak", "double_gaussian"):
            f, kw, _ = make_family(fname, d, rng)
            dom = domain_for(fname, d, rng)
            seen = as_seen(dom, dtype)
            ref = separable_ref(rule, fname, f, seen, n)
            # positive families: sum |w f| = |sum w f|; the oscillatory one's by brute force
            scale = oracle_value(rule, fname, d, n, seen, kw)[1] if fname == "oscillatory" else abs(ref)
            val = RULES[rule]().integrate(f, d, N=n**d, integration_domain=torch.tensor(dom, dtype=dtype, device=dev))
            check(float(val), ref, dtype, scale, parity)


def test_compiled_graphs_own_their_workspaces(dev, parity):
    """Each captured graph owns its workspace: after a larger capture (which
    would regrow a shared one) and after the shared per-stream workspace was
    regrown by eager calls, the first graph still replays the oracle's value."""
    from torchquad_amd import _lib

    integ = tq.Boole()
    comp = integ.get_jit_compiled_integrate(3, N=21**3, integration_domain=[[-1.0, 1.0]] * 3)
    f1 = tq.GaussianPeak([2.0, 3.0, 1.5], [0.1, -0.2, 0.3])
    f2 = tq.ProductPeak([1.0, 2.0, 1.5], [0.0, 0.2, -0.1])
    dom = [[-1.0, 0.5], [-0.5, 1.0], [0.0, 1.0]]
    d32 = torch.tensor(dom, dtype=torch.float32, device=dev)
    first = float(comp(f1, d32))
    float(comp(f2, d32))
    big = integ.get_jit_compiled_integrate(5, N=65**5, integration_domain=[[-1.0, 1.0]] * 5)
    float(big(tq.GaussianPeak([2.0] * 5, [0.0] * 5), torch.tensor([[-1.0, 1.0]] * 5, device=dev)))
    # regrow the shared workspace of this stream
    stream = torch.cuda.current_stream(dev).cuda_stream
    key = (dev.index if dev.index is not None else torch.cuda.current_device(), stream, "main")
    old = _lib.workspace._bufs.get(key)
    _lib.workspace.get((old.numel() if old is not None else 256) * 4, dev)
    torch.cuda.empty_cache()
    cells = comp.__closure__
    graphs = next(c.cell_contents for c in cells if isinstance(c.cell_contents, dict))
    for entry in graphs.values():
        ws = entry[4][0]
        assert all(ws.data_ptr() != b.data_ptr() for b in _lib.workspace._bufs.values())
    ref, scale = oracle_value("boole", "gaussian_peak", 3, 21, as_seen(dom, torch.float32),
                              dict(a=f1.a, u=f1.u))
    again = float(comp(f1, d32))
    assert again == first
    check(again, ref, torch.float32, scale, parity)


@pytest.mark.parametrize("S", [(1,), (7,), (200,), (2, 3)])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_param_batch_fused_vs_oracle(S, dtype, dev, parity):
    """Built-in families holding a parameter batch S (the reference's [N, *S]
    vectorised integrands, base_integrator.py:77-89, tutorial.rst:741-774):
    one batched launch over the grid, every parameter set against the FP64
    oracle (error <= rtol * sum |w f|), the result shaped [*S] (a [N, 1]
    batch is squeezed like a scalar integrand), and equal to the same object
    called as a torch [N, *S] integrand on the point/reduce path."""
    rng = np.random.default_rng(seed_of("param_batch", S, dtype))
    cases = [
        # (rule, dim, N, family, oracle family, parameter builder)
        ("simpson", 1, 10001, "gaussian_peak"),
        ("boole", 3, 17**3, "product_peak"),
        ("gauss_legendre", 2, 24**2, "oscillatory"),
        ("trapezoid", 2, 33**2, "exp_linear"),
    ]
    for rule, d, N, fname in cases:
        a = rng.uniform(1.0, 5.0, S + (d,))
        u = rng.uniform(-0.5, 0.5, S + (d,))
        w = rng.uniform(0.0, 1.0, S)
        if fname == "gaussian_peak":
            f = tq.GaussianPeak(a, u)
            kw = [dict(a=a.reshape(-1, d)[i], u=u.reshape(-1, d)[i]) for i in range(int(np.prod(S)))]
        elif fname == "product_peak":
            f = tq.ProductPeak(a * 0.4, u)
            kw = [dict(a=0.4 * a.reshape(-1, d)[i], u=u.reshape(-1, d)[i]) for i in range(int(np.prod(S)))]
        elif fname == "oscillatory":
            f = tq.Oscillatory(a, w)
            kw = [dict(a=a.reshape(-1, d)[i], c1=float(w.reshape(-1)[i])) for i in range(int(np.prod(S)))]
        else:
            f = tq.ExpLinear(d, a=a * 0.3, c=w)
            kw = [dict(a=0.3 * a.reshape(-1, d)[i], c1=float(w.reshape(-1)[i])) for i in range(int(np.prod(S)))]
        assert f.batch_shape == S
        dom = [[-1.0, 0.5 + 0.1 * k] for k in range(d)]
        domain = torch.tensor(dom, dtype=dtype, device=dev)
        integ = RULES[rule]()
        val = integ.integrate(f, d, N=N, integration_domain=domain)
        n = integ._grid_n(N, d)
        assert val.dtype == dtype
        assert val.shape == (() if S == (1,) else S)
        flat = val.reshape(-1).cpu().numpy()
        for i in range(flat.size):
            ref, scale = oracle_value(rule, fname, d, n, as_seen(dom, dtype), kw[i])
            check(float(flat[i]), ref, dtype, scale, parity)
        # the same object as a plain [N, *S] torch integrand (point writer -> fn -> reduce).
        # In FP32 that path scales Newton-Cotes sums by the reference's h = x_1 - x_0 of the
        # FP32 grid (grid_integrator.py:57-91), whose rounding (up to 4e-4 relative at
        # n = 10001) the fused kernels, with h = (b - a)/(n - 1) in FP64, do not have: the
        # values are compared in FP64.
        user = RULES[rule]().integrate(lambda x: f(x), d, N=N, integration_domain=domain)
        assert user.shape == val.shape
        if dtype == torch.float64:
            np.testing.assert_allclose(user.reshape(-1).cpu().numpy(), flat, rtol=1e-12,
                                       atol=1e-12 * float(np.max(np.abs(flat))))
        # a repeated call reuses the plan and gives the same values
        again = integ.integrate(f, d, N=N, integration_domain=domain)
        assert torch.equal(again, val)
