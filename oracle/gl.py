"""High-precision Gauss-Legendre nodes/weights (oracle).

Newton iteration on the Legendre three-term recurrence in 50-digit mpmath
arithmetic; pins the device generator beyond numpy's leggauss, whose
endpoint weights carry ~1e-10 relative error at n ~ 250.
"""
import mpmath as mp


def node_weight(n, k, dps=50):
    """k-th node (ascending, 0-based) and weight of the n-point rule."""
    with mp.workdps(dps):
        j = n - 1 - k  # descending index for the cosine estimate
        x = mp.cos(mp.pi * (j + mp.mpf(0.75)) / (n + mp.mpf(0.5)))
        for _ in range(100):
            p0, p1 = mp.mpf(1), x
            for m in range(2, n + 1):
                p0, p1 = p1, ((2 * m - 1) * x * p1 - (m - 1) * p0) / m
            dp = n * (x * p1 - p0) / (x * x - 1)
            dx = p1 / dp
            x -= dx
            if abs(dx) < mp.mpf(10) ** (-dps + 5):
                break
        p0, p1 = mp.mpf(1), x
        for m in range(2, n + 1):
            p0, p1 = p1, ((2 * m - 1) * x * p1 - (m - 1) * p0) / m
        dp = n * (x * p1 - p0) / (x * x - 1)
        w = 2 / ((1 - x * x) * dp * dp)
        return float(x), float(w)
