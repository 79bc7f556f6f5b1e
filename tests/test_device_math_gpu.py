"""The device math of the built-in families, point by point.

A one-node Gauss-Legendre rule on [m - 1, m + 1] (node 0, weight 2 up to
its FP64 rounding, axis scale 1) returns 2 f(m) in the FP64 output of the
batched kernel (the FP32 value is accumulated in FP64), so a batch of such
integrals evaluates the device integrand at chosen points.  Checked against FP64 numpy:
- the revolution cosine of the FP32 oscillatory family (range reduction and
  an even polynomial fitted in q^2, stated bound 1.6e-7 absolute);
- exp2 of the FP32 Gaussian peak (hardware v_exp_f32 on the folded exponent);
- the reciprocal of the FP32 product peak (hardware v_rcp_f32).
"""
import math

import numpy as np
import pytest
import torch

import torchquad_amd as tq
from torchquad_amd.batch import Batch

pytestmark = pytest.mark.gpu


def device_values(fn, points):
    items = [{"rule": "gauss_legendre", "fn": fn, "dim": 1, "n": 1, "domain": [[float(m) - 1.0, float(m) + 1.0]]}
             for m in points]
    b = Batch(items, dtype=torch.float32)
    out64 = b.run()[1].cpu().numpy()
    return out64 / 2.0


def test_revolution_cosine_max_error(dev):
    # FP32 points over several revolutions, the quarter-revolution reflection
    # points and their neighbours, and integers
    s = np.float32(np.linspace(-3.0, 3.0, 20001))
    special = np.float32([0.25, -0.25, 0.5, -0.5, 0.75, 1.0, 2.0, -1.0])
    near = np.concatenate([np.nextafter(special, np.float32(np.inf)), np.nextafter(special, np.float32(-np.inf))])
    pts = np.concatenate([s, special, near]).astype(np.float32)
    # c = 2 pi: the FP32 phase is c / (2 pi) = 1 revolution per unit x, w = 0
    f = tq.Oscillatory([2.0 * math.pi], 0.0)
    got = device_values(f, pts)
    ref = np.cos(2.0 * np.pi * pts.astype(np.float64))
    err = np.max(np.abs(got - ref))
    assert err <= 1.6e-7, err


def test_gaussian_exp2_and_product_peak_reciprocal(dev):
    pts = np.float32(np.linspace(-1.5, 1.5, 6001))
    a, u = 2.5, 0.3
    g = device_values(tq.GaussianPeak([a], [u]), pts)
    x = pts.astype(np.float64)
    ref_g = np.exp(-((a * (x - u)) ** 2))
    # v_exp_f32 plus the FP32 rounding of the folded exponent t = a' x + b'
    # (|d(t^2)| ~ 2|t| ulp(t)): relative error grows with the exponent
    t2 = (a * (x - u)) ** 2 / math.log(2.0)
    bound = 4 * 2.0 ** -23 * (1.0 + 2.0 * t2)
    assert np.all(np.abs(g - ref_g) <= bound * ref_g), float(np.max(np.abs(g - ref_g) / ref_g / bound))
    p = device_values(tq.ProductPeak([a], [u]), pts)
    ref_p = 1.0 / (a ** -2 + (x - u) ** 2)
    assert np.max(np.abs(p - ref_p) / ref_p) <= 4 * 2.0 ** -23
