"""Hand-written backward of the Gaussian-peak sweep vs closed forms, central
finite differences and the reference-style autograd restatement."""
import numpy as np
import pytest
import torch

from torchquad_amd.autograd import gaussian_peak_integrals
from oracle import grad as ogr

pytestmark = pytest.mark.gpu


def c5_inputs(configs, B, dtype, dev):
    c = configs["C5"]
    rng = np.random.default_rng(c["seed"])
    a = rng.uniform(1.0, 5.0, (c["count"], 4))[:B]
    u = rng.uniform(-0.5, 0.5, (c["count"], 4))[:B]
    lo = rng.uniform(-1.0, 0.0, (c["count"], 4))[:B]
    dom = np.stack([lo, lo + 1.0], axis=2)
    t = lambda v: torch.tensor(v, dtype=dtype, device=dev, requires_grad=True)  # noqa: E731
    return a, u, lo, t(a), t(u), t(dom)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_forward_and_backward_vs_closed_form(configs, dtype, dev):
    B = 64
    a, u, lo, ta, tu, td = c5_inputs(configs, B, dtype, dev)
    out = gaussian_peak_integrals(ta, tu, td, n=32, rule="gauss_legendre")
    out.sum().backward()
    rtol = 1e-6 if dtype == torch.float32 else 1e-12
    gtol = 2e-5 if dtype == torch.float32 else 1e-10
    for b in range(B):
        I, da, du, dlo, dhi = ogr.closed_form(a[b], u[b], lo[b], lo[b] + 1.0)
        assert abs(float(out[b]) - I) <= rtol * I, (b, float(out[b]), I)
        scale = max(np.max(np.abs(da)), I)
        np.testing.assert_allclose(ta.grad[b].cpu().numpy(), da, atol=gtol * scale)
        np.testing.assert_allclose(tu.grad[b].cpu().numpy(), du, atol=gtol * scale)
        np.testing.assert_allclose(td.grad[b, :, 0].cpu().numpy(), dlo, atol=gtol * scale)
        np.testing.assert_allclose(td.grad[b, :, 1].cpu().numpy(), dhi, atol=gtol * scale)


def test_backward_vs_reference_autograd(configs, dev):
    B = 3
    a, u, lo, ta, tu, td = c5_inputs(configs, B, torch.float64, dev)
    out = gaussian_peak_integrals(ta, tu, td, n=32)
    out.sum().backward()
    for b in range(B):
        I, da, du, dlo, dhi = ogr.reference_autograd(a[b], u[b], lo[b], lo[b] + 1.0, 32)
        assert abs(float(out[b]) - I) <= 1e-12 * I
        np.testing.assert_allclose(ta.grad[b].cpu().numpy(), da, rtol=1e-9, atol=1e-13)
        np.testing.assert_allclose(tu.grad[b].cpu().numpy(), du, rtol=1e-9, atol=1e-13)
        np.testing.assert_allclose(td.grad[b, :, 0].cpu().numpy(), dlo, rtol=1e-9, atol=1e-13)
        np.testing.assert_allclose(td.grad[b, :, 1].cpu().numpy(), dhi, rtol=1e-9, atol=1e-13)


@pytest.mark.parametrize("rule,n", [("gauss_legendre", 9), ("trapezoid", 21), ("simpson", 21), ("boole", 17)])
def test_backward_vs_central_differences(rule, n, dev):
    rng = np.random.default_rng(7)
    B, d, h = 4, 3, 1e-3
    a = torch.tensor(rng.uniform(1, 3, (B, d)), device=dev, requires_grad=True)
    u = torch.tensor(rng.uniform(-0.3, 0.3, (B, d)), device=dev, requires_grad=True)
    lo = rng.uniform(-1, -0.2, (B, d))
    dom = torch.tensor(np.stack([lo, lo + rng.uniform(0.8, 1.5, (B, d))], 2), device=dev, requires_grad=True)
    out = gaussian_peak_integrals(a, u, dom, n=n, rule=rule)
    out.sum().backward()

    def f(a_, u_, d_):
        with torch.no_grad():
            return gaussian_peak_integrals(a_, u_, d_, n=n, rule=rule)

    for name, t, g in (("a", a, a.grad), ("u", u, u.grad), ("dom", dom, dom.grad)):
        base = [a.detach().clone(), u.detach().clone(), dom.detach().clone()]
        idx = {"a": 0, "u": 1, "dom": 2}[name]
        flat = base[idx].view(-1)
        fd = torch.zeros_like(flat)
        for i in range(flat.numel()):
            plus = [v.clone() for v in base]
            minus = [v.clone() for v in base]
            plus[idx].view(-1)[i] += h
            minus[idx].view(-1)[i] -= h
            diff = f(*plus) - f(*minus)
            fd[i] = diff.sum() / (2 * h)
        np.testing.assert_allclose(g.view(-1).cpu().numpy(), fd.cpu().numpy(), rtol=1e-5, atol=1e-7)


def test_n_from_N_and_validation(dev):
    a = torch.ones(2, 2, device=dev, dtype=torch.float64)
    u = torch.zeros(2, 2, device=dev, dtype=torch.float64)
    dom = torch.tensor([[[-1.0, 1.0]] * 2] * 2, device=dev, dtype=torch.float64)
    v1 = gaussian_peak_integrals(a, u, dom, N=100, rule="simpson")
    v2 = gaussian_peak_integrals(a, u, dom, n=9, rule="simpson")
    assert torch.equal(v1, v2)
    with pytest.raises(ValueError):
        gaussian_peak_integrals(a, u, dom[:, :1], n=9)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_long_loop_axis_without_table(dtype, dev):
    """n > 1024 on the loop axis: the backward computes its loop terms on the fly."""
    a = np.array([[2.0, 1.5], [3.0, 2.5]])
    u = np.array([[0.1, -0.2], [0.3, 0.0]])
    lo = np.array([[-1.0, -0.5], [-0.8, -1.0]])
    hi = lo + 1.3
    t = lambda v: torch.tensor(v, dtype=dtype, device=dev, requires_grad=True)  # noqa: E731
    ta, tu, td = t(a), t(u), t(np.stack([lo, hi], 2))
    out = gaussian_peak_integrals(ta, tu, td, n=1500, rule="gauss_legendre")
    out.sum().backward()
    tol = 2e-5 if dtype == torch.float32 else 1e-10
    for b in range(2):
        I, da, du, dlo, dhi = ogr.closed_form(a[b], u[b], lo[b], hi[b])
        assert abs(float(out[b]) - I) <= (1e-6 if dtype == torch.float32 else 1e-12) * I
        scale = max(np.max(np.abs(da)), I)
        np.testing.assert_allclose(ta.grad[b].cpu().numpy(), da, atol=tol * scale)
        np.testing.assert_allclose(tu.grad[b].cpu().numpy(), du, atol=tol * scale)
        np.testing.assert_allclose(td.grad[b, :, 0].cpu().numpy(), dlo, atol=tol * scale)
        np.testing.assert_allclose(td.grad[b, :, 1].cpu().numpy(), dhi, atol=tol * scale)
