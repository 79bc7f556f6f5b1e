"""Batched form on the GPU vs the oracle, the plain path and itself."""
import math
import warnings

import numpy as np
import pytest
import torch

import torchquad_amd as tq
from torchquad_amd.batch import Batch, n_per_axis
from torchquad_amd.workloads import genz_batch
from oracle import families as fam
from oracle import grid as og
from oracle import native

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _quiet():
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        yield


def oracle_value(it, n):
    dim, rule, dom = it["dim"], it["rule"], it["domain"]
    a, u, w = it["a"], it["u"], it["w"]
    if rule == "gauss_legendre" and n > 1000:
        lo, hi = [0.0] * dim, [1.0] * dim
        return {"gaussian_peak": lambda: fam.gaussian_peak_integral(a, u, lo, hi),
                "oscillatory": lambda: fam.oscillatory_integral(a, w, lo, hi),
                "product_peak": lambda: fam.product_peak_integral(a, u, lo, hi)}[it["family"]]()
    if it["family"] == "gaussian_peak":
        return og.separable_gaussian_peak(rule, a, u, dom, n)
    if it["family"] == "oscillatory":
        return og.separable_oscillatory(rule, a, w, dom, n)
    return og.separable_product(rule, [lambda x, ak=ak, uk=uk: 1.0 / (ak**-2.0 + (x - uk) ** 2)
                                       for ak, uk in zip(a, u)], dom, n)


def scale_of(it, n, ref):
    """sum |w f|: |ref| for the positive families; for the oscillatory one the
    brute-force FP64 sum, or for one-axis rules too long for leggauss the
    integral of |f| by a 2e6-point trapezoid sum (a scale, not a reference)."""
    if it["family"] != "oscillatory":
        return abs(ref)
    rule = it["rule"]
    if rule == "gauss_legendre" and n > 1000:
        x = np.linspace(0.0, 1.0, 2_000_001)
        y = np.abs(np.cos(2 * np.pi * it["w"] + it["a"][0] * x))
        return float(np.sum(y[1:] + y[:-1]) * 0.5 / 2_000_000)
    return native.grid_sum(rule, "oscillatory", it["dim"], n, it["domain"], c1=it["w"], a=it["a"], with_abs=True)[1]


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_batch_vs_oracle(dtype, dev, parity):
    """300 mixed integrals against FP64 separable sums (closed forms for
    Gauss-Legendre beyond n = 1000, where the rule is exact to FP64): error
    <= rtol * sum |w f|, rtol 1e-6 (FP32) or 1e-12 (FP64)."""
    items = genz_batch(300, seed=3)
    b = Batch(items, dtype=dtype)
    _, out64 = b.run()
    res = out64.cpu().numpy()
    tol = 1e-6 if dtype == torch.float32 else 1e-12
    for i, it in enumerate(items):
        rule = tq.batch.RULE_NAMES[it["rule"]]
        n = n_per_axis(rule, it["dim"], it["N"])
        ref = oracle_value(it, n)
        scale = scale_of(it, n, ref)
        parity.record(abs(res[i] - ref) / scale, tol)
        assert abs(res[i] - ref) <= tol * scale, (i, it["rule"], it["family"], it["dim"], n, res[i], ref)


def test_batch_matches_plain_path_and_is_deterministic(dev):
    items = genz_batch(64, seed=9)
    b = Batch(items, dtype=torch.float32)
    out, out64 = b.run()
    first = out64.clone()
    b.run()
    assert torch.equal(first, b.out64)
    rules = {"trapezoid": tq.Trapezoid, "simpson": tq.Simpson, "gauss_legendre": tq.GaussLegendre}
    for i, it in enumerate(items[:24]):
        dom = torch.tensor(it["domain"], dtype=torch.float32, device=dev)
        plain = float(rules[it["rule"]]().integrate(it["fn"], it["dim"], N=it["N"], integration_domain=dom))
        rule = tq.batch.RULE_NAMES[it["rule"]]
        n = n_per_axis(rule, it["dim"], it["N"])
        scale = scale_of(it, n, oracle_value(it, n))
        assert abs(plain - float(out[i])) <= 2e-6 * scale, (i, plain, float(out[i]))
    # a different tiling gives the same values up to FP32 partial rounding
    b.retile(tile_points=2048)
    b.run()
    np.testing.assert_allclose(b.out64.cpu().numpy(), first.cpu().numpy(), rtol=2e-7, atol=1e-9)


def test_fp32_gl_pairs_vs_fp64_tables(dev):
    """FP32 batches read one-axis Gauss-Legendre rows from FP32 (node, weight) pairs; without the
    pairs (NULL) the same kernel reads the FP64 tables.  Both against the oracle at the FP32 bar,
    odd and even n (the middle node of odd n is counted once)."""
    rng = np.random.default_rng(21)
    items = []
    for n in (128, 129, 255, 256, 1001, 4096, 65537, 200_001):
        a, u = rng.uniform(1.0, 8.0, 1), rng.uniform(0.2, 0.8, 1)
        items.append({"rule": "gauss_legendre", "fn": tq.GaussianPeak(a, u), "dim": 1, "n": n,
                      "domain": [[0.0, 1.0]], "family": "gaussian_peak", "a": a, "u": u, "w": None})
    b = Batch(items, dtype=torch.float32)
    assert b.gl_p32 is not None
    with_pairs = b.run()[1].clone()
    b.gl_p32 = None
    fp64_tabs = b.run()[1].clone()
    for i, it in enumerate(items):
        if it["n"] <= 4096:
            ref = og.separable_gaussian_peak("gauss_legendre", it["a"], it["u"], it["domain"], it["n"])
        else:  # the rule is exact to FP64 rounding at these n: the analytic integral
            ref = fam.gaussian_peak_integral(it["a"], it["u"], [0.0], [1.0])
        for v in (float(with_pairs[i]), float(fp64_tabs[i])):
            assert abs(v - ref) <= 1e-6 * abs(ref), (it["n"], v, ref)


def test_gl_tables_shorter_than_an_item_give_nan(dev):
    """A Gauss-Legendre item whose n nodes at gl_offset do not all lie in the
    gl_len entries passed to tq_batch_run reads no table and gives NaN; the
    other items are unchanged.  (The tables really hold 150 entries, so no
    read leaves the allocation whatever the kernel does.)"""
    f = tq.GaussianPeak([3.0], [0.4])
    items = [{"rule": "gauss_legendre", "fn": f, "dim": 1, "n": 100, "domain": [[0.0, 1.0]]},
             {"rule": "gauss_legendre", "fn": f, "dim": 1, "n": 50, "domain": [[0.0, 1.0]]},
             {"rule": "trapezoid", "fn": f, "dim": 1, "n": 333, "domain": [[0.0, 1.0]]}]
    for dtype in (torch.float32, torch.float64):
        b = Batch(items, dtype=dtype)
        assert b.gl_x.numel() == 150
        full = b.run()[1].clone()
        assert torch.isfinite(full).all()
        b.gl_x, b.gl_w = b.gl_x[:120], b.gl_w[:120]  # the second table (offset 100, n 50) ends at 150
        short = b.run()[1].clone()
        assert float(short[0]) == float(full[0]) and float(short[2]) == float(full[2])
        assert math.isnan(float(short[1]))


def test_batch_graph_replay(dev):
    items = genz_batch(128, seed=4)
    b = Batch(items, dtype=torch.float32)
    _, ref = b.run()
    ref = ref.clone()
    g = b.graph()
    for _ in range(3):
        b.out64.zero_()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(b.out64, ref)
        # eager work between replays
        torch.ones(1000, device=dev).sum().item()


def test_autotune_picks_a_candidate(dev):
    from torchquad_amd import autotune

    b = Batch(genz_batch(200, seed=5), dtype=torch.float32)
    _, ref = b.run()
    ref = ref.clone()
    tp, table = autotune.autotune(b, tiles=(2048, 8192), reps=2, use_cache=False)
    assert tp in {t for t, _ in table}
    b.run()
    # the tuned geometry gives exactly what a fresh batch with that geometry gives
    same = Batch(genz_batch(200, seed=5), dtype=torch.float32, tile_points=tp)
    assert torch.equal(b.out64, same.run()[1])
    # another tile size only regroups FP32 partial sums: FP32 bar
    np.testing.assert_allclose(b.out64.cpu().numpy(), ref.cpu().numpy(), rtol=1e-6, atol=1e-9)


def test_gl_generator_large_n(dev):
    from oracle import gl as ogl
    from torchquad_amd.gl_nodes import gauss_legendre

    for n in (1000, 5000):
        x, w = gauss_legendre(n, dev)
        xs, ws = x.cpu().numpy(), w.cpu().numpy()
        for k in (0, 1, 15, 16, 31, 40, 200, n // 2 - 1):  # 15 | 16: boundary / interior switch
            mx, mw = ogl.node_weight(n, k)
            assert abs(xs[k] - mx) <= 2e-16 * max(1.0, n / 1000), (n, k, xs[k] - mx)
            assert abs(ws[k] - mw) <= 2e-12 * mw, (n, k, (ws[k] - mw) / mw)
    for n in (200_000, 1_000_000):
        x, w = gauss_legendre(n, dev)
        assert abs(float(w.sum()) - 2.0) < 1e-12
        assert abs(float(torch.sum(w * x * x)) - 2.0 / 3.0) < 1e-12
        assert float((x[1:] - x[:-1]).min()) > 0
        assert torch.allclose(x, -x.flip(0), atol=1e-16, rtol=0)
