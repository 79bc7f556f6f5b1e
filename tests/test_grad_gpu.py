"""Hand-written backward of the parameter sweeps (Gaussian peak, oscillatory,
product peak) vs closed forms, central finite differences and the
reference-style autograd restatement."""
import zlib

import numpy as np
import pytest
import torch

from torchquad_amd.autograd import gaussian_peak_integrals, oscillatory_integrals, product_peak_integrals
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


FAMILY_SWEEPS = {"oscillatory": oscillatory_integrals, "product_peak": product_peak_integrals}


def family_inputs(family, B, d, seed):
    rng = np.random.default_rng(seed)
    if family == "oscillatory":
        p1 = rng.uniform(0.5, 3.0, (B, d)) * rng.choice([-1.0, 1.0], (B, d))
        p2 = rng.uniform(0.0, 1.0, B)
    else:
        p1 = rng.uniform(1.0, 5.0, (B, d))
        p2 = rng.uniform(-0.5, 0.5, (B, d))
    lo = rng.uniform(-1.0, 0.0, (B, d))
    hi = lo + rng.uniform(0.6, 1.4, (B, d))
    return p1, p2, lo, hi


def _tensors(dtype, dev, *arrays):
    return [torch.tensor(v, dtype=dtype, device=dev, requires_grad=True) for v in arrays]


# one-axis split rows (d=1, n >= 128), lane-only (d=1, n < 128), loop tables,
# a loop axis without a table (n > 1024), and every rule
FAMILY_SHAPES = [("gauss_legendre", 32, 4), ("boole", 17, 3), ("simpson", 21, 2), ("trapezoid", 65, 1),
                 ("gauss_legendre", 300, 1), ("gauss_legendre", 1100, 2)]


@pytest.mark.parametrize("family", ["oscillatory", "product_peak"])
@pytest.mark.parametrize("rule,n,d", FAMILY_SHAPES)
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_family_sweep_vs_reference_autograd(family, rule, n, d, dtype, dev, parity):
    """Forward and hand-written backward against torch autograd through the
    reference's materialised grid (FP64 CPU).  Bars: values 1e-12 (FP64) /
    1e-6 (FP32) of the integral for the positive product peak and of the box
    volume, which bounds sum |w f|, for the oscillatory family; gradients
    1e-9 / 2e-5 of the integral's largest gradient component."""
    B = 3
    p1, p2, lo, hi = family_inputs(family, B, d, zlib.crc32(f"{family}/{rule}/{n}/{d}".encode()))
    t1, t2, td = _tensors(dtype, dev, p1, p2, np.stack([lo, hi], 2))
    out = FAMILY_SWEEPS[family](t1, t2, td, n=n, rule=rule)
    out.sum().backward()
    f64 = dtype == torch.float64
    vtol, gtol = (1e-12, 1e-9) if f64 else (1e-6, 2e-5)
    # FP32 inputs as the kernel sees them
    q1, q2, qlo, qhi = (np.asarray(torch.tensor(v, dtype=dtype).double()) for v in (p1, p2, lo, hi))
    for b in range(B):
        I, g1, g2, glo, ghi = ogr.reference_autograd_family(family, q1[b], q2[b], qlo[b], qhi[b], n, rule)
        scale = abs(I) if family == "product_peak" else float(np.prod(qhi[b] - qlo[b]))
        err = abs(float(out[b]) - I) / scale
        parity.record(err, vtol)
        assert err <= vtol, (b, float(out[b]), I)
        grads = [(t1.grad[b], g1), (t2.grad[b], g2), (td.grad[b, :, 0], glo), (td.grad[b, :, 1], ghi)]
        gscale = max(float(np.max(np.abs(np.atleast_1d(r)))) for _, r in grads)
        for got, ref in grads:
            gerr = float(np.max(np.abs(np.atleast_1d(got.detach().cpu().numpy()) - np.atleast_1d(ref)))) / gscale
            parity.record(gerr, gtol)
            assert gerr <= gtol, (b, got, ref)


@pytest.mark.parametrize("family", ["oscillatory", "product_peak"])
def test_family_backward_vs_central_differences(family, dev):
    rule, n, d, B, h = "boole", 9, 3, 2, 1e-5  # FD truncation ~h^2 f''' below the 1e-6 bar
    p1, p2, lo, hi = family_inputs(family, B, d, 5)
    t1, t2, td = _tensors(torch.float64, dev, p1, p2, np.stack([lo, hi], 2))
    fn = FAMILY_SWEEPS[family]
    out = fn(t1, t2, td, n=n, rule=rule)
    out.sum().backward()
    base = [t1.detach().clone(), t2.detach().clone(), td.detach().clone()]
    for idx, g in enumerate((t1.grad, t2.grad, td.grad)):
        flat = base[idx].view(-1)
        fd = torch.zeros_like(flat)
        for i in range(flat.numel()):
            plus = [v.clone() for v in base]
            minus = [v.clone() for v in base]
            plus[idx].view(-1)[i] += h
            minus[idx].view(-1)[i] -= h
            with torch.no_grad():
                fd[i] = (fn(*plus, n=n, rule=rule) - fn(*minus, n=n, rule=rule)).sum() / (2 * h)
        np.testing.assert_allclose(g.view(-1).cpu().numpy(), fd.cpu().numpy(), rtol=1e-6, atol=1e-9)


def test_family_sweeps_vs_closed_forms(dev):
    """FP64 Gauss-Legendre n=40 is exact to rounding for these smooth
    integrands: the sweep values equal the closed-form integrals."""
    from oracle import families as fam

    for family, closed in (("oscillatory", fam.oscillatory_integral), ("product_peak", fam.product_peak_integral)):
        p1, p2, lo, hi = family_inputs(family, 4, 3, 11)
        if family == "product_peak":
            p1 = p1 / 2.5  # widths that 40 nodes resolve to rounding
        t1, t2, td = _tensors(torch.float64, dev, p1, p2, np.stack([lo, hi], 2))
        out = FAMILY_SWEEPS[family](t1, t2, td, n=40)
        for b in range(4):
            ref = closed(p1[b], p2[b], lo[b], hi[b])
            assert abs(float(out[b]) - ref) <= 1e-11 * max(abs(ref), float(np.prod(hi[b] - lo[b]))), (family, b)
