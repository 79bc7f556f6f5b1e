"""FP64 numpy integrand families and closed-form box integrals (oracle).

Pinned by: closed forms vs high-order quadrature (scipy) on small boxes, and
agreement of the separable product sums with the materialised collapse
(tests/test_oracle.py).
"""
import math

import numpy as np
from scipy.special import erf


def gaussian_peak(x, a, u):
    return np.exp(-np.sum((a * (x - u)) ** 2, axis=-1))


def oscillatory(x, c, w):
    return np.cos(2.0 * math.pi * w + x @ c)


def product_peak(x, a, u):
    return np.prod(1.0 / (a**-2.0 + (x - u) ** 2), axis=-1)


def exp_linear(x, a, c=0.0):
    return np.exp(c + x @ a)


def sin_sum(x, a, u):
    return np.sum(np.sin(a * x + u), axis=-1)


def double_gaussian(x, width=0.1, centers=(1.0 / 3.0, 2.0 / 3.0)):
    d = x.shape[-1]
    norm = 0.5 * (width * math.sqrt(math.pi)) ** (-d)
    inv = 1.0 / (width * width)
    return norm * (
        np.exp(-inv * np.sum((x - centers[0]) ** 2, axis=-1)) + np.exp(-inv * np.sum((x - centers[1]) ** 2, axis=-1))
    )


# ---- closed forms over the box prod [lo_k, hi_k] ---------------------------
def gaussian_peak_integral(a, u, lo, hi):
    a, u, lo, hi = map(np.asarray, (a, u, lo, hi))
    per = math.sqrt(math.pi) / (2.0 * a) * (erf(a * (hi - u)) - erf(a * (lo - u)))
    return float(np.prod(per))


def oscillatory_integral(c, w, lo, hi):
    """Re[ e^{i 2 pi w} prod_k (e^{i c_k hi_k} - e^{i c_k lo_k}) / (i c_k) ] (c_k != 0)."""
    z = complex(math.cos(2 * math.pi * w), math.sin(2 * math.pi * w))
    for ck, l, h in zip(c, lo, hi):
        if ck == 0.0:
            z *= h - l
        else:
            z *= (complex(math.cos(ck * h), math.sin(ck * h)) - complex(math.cos(ck * l), math.sin(ck * l))) / complex(0.0, ck)
    return z.real


def product_peak_integral(a, u, lo, hi):
    """prod_k a_k [atan(a_k (hi-u)) - atan(a_k (lo-u))]."""
    val = 1.0
    for ak, uk, l, h in zip(a, u, lo, hi):
        val *= ak * (math.atan(ak * (h - uk)) - math.atan(ak * (l - uk)))
    return val


def exp_linear_integral(a, c, lo, hi):
    val = math.exp(c)
    for ak, l, h in zip(a, lo, hi):
        val *= (math.exp(ak * h) - math.exp(ak * l)) / ak if ak != 0 else (h - l)
    return val


def sin_sum_integral(a, u, lo, hi):
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    vol = float(np.prod(hi - lo))
    total = 0.0
    for k, (ak, uk) in enumerate(zip(a, u)):
        one = (math.cos(ak * lo[k] + uk) - math.cos(ak * hi[k] + uk)) / ak
        total += one * vol / (hi[k] - lo[k])
    return total


def double_gaussian_integral(d, lo, hi, width=0.1, centers=(1.0 / 3.0, 2.0 / 3.0)):
    norm = 0.5 * (width * math.sqrt(math.pi)) ** (-d)
    total = 0.0
    for c in centers:
        p = 1.0
        for l, h in zip(lo, hi):
            p *= 0.5 * width * math.sqrt(math.pi) * (erf((h - c) / width) - erf((l - c) / width))
        total += p
    return norm * total
