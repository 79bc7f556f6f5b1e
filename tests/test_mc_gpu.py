"""Monte Carlo and the Philox stream on the GPU vs the oracle."""
import math
import warnings

import numpy as np
import pytest
import torch

import torchquad_amd as tq
from torchquad_amd import _ops
from oracle import families as fam
from oracle import native

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _quiet():
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        yield


@pytest.mark.parametrize("cols", [1, 3, 4, 8, 9])
def test_uniform_stream_bit_exact(cols, dev):
    for dtype, npd in ((torch.float32, np.float32), (torch.float64, np.float64)):
        for seed, off in ((0, 0), (12345678901234, 2**33 + 5)):
            g = _ops.uniform(seed, off, 1000, cols, dtype, dev).cpu().numpy()
            r = native.uniform(seed, off, 1000, cols, npd)
            assert np.array_equal(g, r)


def test_rng_class_reproducible(dev):
    a = tq.RNG(seed=3).uniform([100, 4], dtype=torch.float32)
    r = tq.RNG(seed=3)
    b1 = r.uniform([50, 4], dtype=torch.float32)
    b2 = r.uniform([50, 4], dtype=torch.float32)
    assert torch.equal(a, torch.cat([b1, b2]))
    assert not torch.equal(b1, b2)


@pytest.mark.parametrize("d", [3, 5, 8, 12])  # the kernel's dimension bounds 4, 8 and 16
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("fname", ["gaussian_peak", "oscillatory", "product_peak", "exp_linear", "constant"])
def test_fused_mc_vs_same_stream_oracle(dtype, fname, d, dev, parity):
    """Same Philox stream as the oracle: |value - oracle| <= rtol * volume *
    sum |f| / N (rtol 1e-6 FP32, 1e-12 FP64; plain relative for positive f)."""
    rng = np.random.default_rng(5)
    N, seed = 200_000, 77
    a, u, c = rng.uniform(1, 3, d), rng.uniform(-0.5, 0.5, d), rng.uniform(0, 2, d)
    dom = [[-0.5, 1.0]] * d
    f, kw = {
        "gaussian_peak": (tq.GaussianPeak(a, u), dict(a=a, u=u)),
        "oscillatory": (tq.Oscillatory(c, 0.2), dict(c1=0.2, a=c)),
        "product_peak": (tq.ProductPeak(a, u), dict(a=a, u=u)),
        "exp_linear": (tq.ExpLinear(d, a=c), dict(a=c)),
        "constant": (tq.Constant(d, 1.0), dict()),
    }[fname]
    s1, _, sa = native.mc_sums("float32" if dtype == torch.float32 else "float64", fname, d, N, seed, 0, dom,
                               with_abs=True, **kw)
    vol = 1.5**d
    ref = vol * s1 / N
    domain = torch.tensor(dom, dtype=dtype, device=dev)
    val = float(tq.MonteCarlo().integrate(f, d, N=N, integration_domain=domain, seed=seed))
    rtol = 1e-6 if dtype == torch.float32 else 1e-12
    scale = vol * sa / N
    tol = rtol * scale
    parity.record(abs(val - ref) / scale, rtol)
    assert abs(val - ref) <= tol, (val, ref)
    # the user-callable path sees the same points
    user = float(tq.MonteCarlo().integrate(lambda x: f(x), d, N=N, integration_domain=domain, seed=seed))
    assert abs(user - ref) <= tol


def test_constant_integrand_exact(dev):
    dom = torch.tensor([[0.0, 3.0], [1.0, 2.0]], dtype=torch.float32, device=dev)
    assert float(tq.MonteCarlo().integrate(tq.Constant(2, 2.0), 2, N=10007, integration_domain=dom, seed=0)) == 6.0
    assert float(tq.MonteCarlo().integrate(lambda x: x[:, 0] * 0 + 2.0, 2, N=10007, integration_domain=dom,
                                           seed=0)) == 6.0


def test_baseline_mc_config_full_size(configs, dev, parity):
    """The Monte Carlo workload at full size: d = 8, N = 1e8, FP32 oscillatory.
    Within 3 reported standard errors of the analytic value; the sums over
    [0, N) equal those over [0, N/2) plus [N/2, N) (the stream splits by
    counter offset); the last 2e5 samples of the stream match the oracle."""
    c = configs["C2b"]
    N = 100_000_000
    mc = tq.MonteCarlo()
    f = tq.Oscillatory(c["c"], c["w"])
    dom = torch.tensor(c["domain"], dtype=torch.float32, device=dev)
    val = float(mc.integrate(f, 8, N=N, integration_domain=dom, seed=0))
    se = mc.std_error()
    assert 0 < se < 1e-4
    assert abs(val - c["analytic"]) <= 3 * se, (val, c["analytic"], se)
    whole = torch.zeros(3, dtype=torch.float64, device=dev)
    halves = torch.zeros(2, 3, dtype=torch.float64, device=dev)
    _ops.mc_integrate(f, 8, N, dom, 0, 0, stats=whole)
    _ops.mc_integrate(f, 8, N // 2, dom, 0, 0, stats=halves[0])
    _ops.mc_integrate(f, 8, N - N // 2, dom, 0, N // 2, stats=halves[1])
    w, h = whole.cpu().numpy(), halves.sum(0).cpu().numpy()
    assert w[2] == N and h[2] == N
    # each thread sums up to 32 samples in FP32 before its FP64 accumulation, and
    # the split moves those groups: about 1e-7 relative per group, random over
    # ~3e6 groups
    np.testing.assert_allclose(h[:2], w[:2], rtol=1e-9)
    tail = 200_000
    part = torch.zeros(3, dtype=torch.float64, device=dev)
    _ops.mc_integrate(f, 8, tail, dom, 0, N - tail, stats=part)
    s1, s2, sa = native.mc_sums("float32", "oscillatory", 8, tail, 0, N - tail, c["domain"], c1=c["w"], a=c["c"],
                                with_abs=True)
    got = part.cpu().numpy()
    parity.record(abs(got[0] - s1) / sa, 1e-6)
    assert abs(got[0] - s1) <= 1e-6 * sa, (got[0], s1)
    assert abs(got[1] - s2) <= 1e-6 * s2


def test_mc_gradients_reference_answers(ref_vectors, dev):
    G = ref_vectors["gradients"]
    mc = tq.MonteCarlo()
    dom = torch.tensor(G["poly_domain"]["domain"], dtype=torch.float64, device=dev, requires_grad=True)
    val = mc.integrate(lambda x: 1.0 + 2.0 * x[:, 0] + 3.0 * x[:, 1] ** 2, 2, N=99997, integration_domain=dom, seed=0)
    val.backward()
    assert abs(float(val) - 12.0) < 8e-2
    np.testing.assert_array_less(np.abs(dom.grad.cpu().numpy() - np.array(G["poly_domain"]["grad"])), 0.1)
    param = torch.tensor([1.0, 2.0, 3.0], dtype=torch.float64, device=dev, requires_grad=True)
    val = mc.integrate(lambda x: param[0] + param[1] * x[:, 0] + param[2] * x[:, 1] ** 2, 2, N=99997,
                       integration_domain=[[0.0, 1.0], [0.0, 2.0]], seed=0)
    val.backward()
    np.testing.assert_array_less(np.abs(param.grad.cpu().numpy() - np.array([2.0, 1.0, 8.0 / 3.0])), 5e-2)


def test_split_api_kwargs(dev):
    # monte_carlo_test.py:136-163: positional, keyword and mixed calls agree
    dom = torch.tensor([[0.0, 1.0]], device=dev)
    mc = tq.MonteCarlo()
    pts = mc.calculate_sample_points(1000, dom, seed=42)
    assert pts.shape == (1000, 1) and pts.device == dom.device
    vals, _ = mc.evaluate_integrand(lambda x: torch.exp(x[:, 0]), pts)
    r1 = mc.calculate_result(vals, dom)
    r2 = mc.calculate_result(function_values=vals, integration_domain=dom)
    r3 = mc.calculate_result(vals, integration_domain=dom)
    assert torch.allclose(r1, r2, rtol=1e-10) and torch.allclose(r1, r3, rtol=1e-10)
    # the split path and integrate() see the same points
    whole = tq.MonteCarlo().integrate(lambda x: torch.exp(x[:, 0]), 1, N=1000, integration_domain=dom, seed=42)
    assert float(whole) == float(r1)


def test_compiled_integrate_graph_replays(dev):
    """MonteCarlo.get_jit_compiled_integrate (monte_carlo.py:108-225): every call
    draws new points from one stream started at the seed; for a built-in
    integrand a captured graph replays the fused step.  Each replay equals the
    eager kernel at the same stream position (with eager work and host
    synchronisations between replays); user callables share the stream."""
    d, N, seed = 4, 200_003, 12
    f = tq.GaussianPeak([2.0, 3.0, 1.5, 1.0], [0.1, -0.2, 0.3, 0.0])
    mc = tq.MonteCarlo()
    comp = mc.get_jit_compiled_integrate(d, N=N, integration_domain=[[-1.0, 1.0]] * d, seed=seed)
    rng = np.random.default_rng(3)
    seen = set()
    for rep in range(5):
        lo = rng.uniform(-1.0, 0.0, d)
        dom = torch.tensor(np.stack([lo, lo + rng.uniform(0.5, 2.0, d)], 1), dtype=torch.float32, device=dev)
        got = comp(f, dom)
        stats = torch.zeros(3, dtype=torch.float64, device=dev)
        ref = _ops.mc_integrate(f, d, N, dom, seed, rep * N, stats=stats)
        junk = torch.randn(4096, device=dev).sum()
        torch.cuda.synchronize()
        assert got.shape == () and float(got) == float(ref), (rep, float(got), float(ref))
        se = mc.std_error()
        s1, s2, n = stats.tolist()
        assert n == N and se == pytest.approx(math.sqrt(max(s2 / n - (s1 / n) ** 2, 0.0) / (n - 1)), rel=1e-12)
        seen.add(float(got))
        assert float(junk) == float(junk)
    # the same domain twice gives different samples, hence (almost surely) different values
    dom = torch.tensor([[0.0, 1.0]] * d, dtype=torch.float32, device=dev)
    a, b = float(comp(f, dom)), float(comp(f, dom))
    assert a != b
    # a user callable continues the same stream (position 7 N)
    user = float(comp(lambda x: f(x), dom))
    ref = float(_ops.mc_integrate(f, d, N, dom, seed, 7 * N))
    assert abs(user - ref) <= 1e-6 * abs(ref)


@pytest.mark.parametrize("S", [(1,), (5,), (2, 3)])
def test_param_batch_mc_same_points(S, dev, parity):
    """A parameter batch S on Monte Carlo: every set integrated over the same
    samples (as an [N, *S] integrand sees one point set), each equal to the
    scalar integrand's fused result on that stream, shaped [*S]."""
    d, N, seed = 3, 100_000, 5
    rng = np.random.default_rng(seed_of_mc(S))
    a = rng.uniform(1.0, 3.0, S + (d,))
    u = rng.uniform(-0.5, 0.5, S + (d,))
    f = tq.GaussianPeak(a, u)
    dom = torch.tensor([[-0.5, 1.0]] * d, dtype=torch.float32, device=dev)
    mc = tq.MonteCarlo()
    val = mc.integrate(f, d, N=N, integration_domain=dom, seed=seed)
    assert val.shape == (() if S == (1,) else S)
    flat = val.reshape(-1).cpu().numpy()
    errs = np.asarray(mc.std_error()).reshape(-1)
    for i in range(flat.size):
        one = tq.GaussianPeak(a.reshape(-1, d)[i], u.reshape(-1, d)[i])
        m1 = tq.MonteCarlo()
        ref = float(m1.integrate(one, d, N=N, integration_domain=dom, seed=seed))
        assert flat[i] == ref
        assert errs[i] == pytest.approx(m1.std_error(), rel=1e-12)
        s1, _, sa = native.mc_sums("float32", "gaussian_peak", d, N, seed, 0, [[-0.5, 1.0]] * d,
                                   a=a.reshape(-1, d)[i], u=u.reshape(-1, d)[i], with_abs=True)
        vol = 1.5**d
        parity.record(abs(ref - vol * s1 / N) / (vol * sa / N), 1e-6)
        assert abs(ref - vol * s1 / N) <= 1e-6 * vol * sa / N


def seed_of_mc(S):
    import zlib

    return zlib.crc32(repr(S).encode())
