"""ctypes binding of liboracle.so (oracle.c) — test infrastructure only."""
import ctypes
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_lib = None


def lib():
    global _lib
    if _lib is None:
        path = _HERE / "liboracle.so"
        if not path.exists():
            from .build import build

            build()
        L = ctypes.CDLL(str(path))
        D, P, I, I64, U64 = ctypes.c_double, ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_uint64
        L.oracle_philox4x32_10.argtypes = [P, P, P]
        L.oracle_uniform_f32.argtypes = [U64, U64, I64, ctypes.c_int32, P]
        L.oracle_uniform_f64.argtypes = [U64, U64, I64, ctypes.c_int32, P]
        L.oracle_grid_sum.restype = D
        L.oracle_grid_sum.argtypes = [I, I, I, I, P, D, D, D, D, P, P, P, P, P]
        L.oracle_mc_sum.argtypes = [I, I, I, I64, U64, U64, P, D, D, D, D, P, P, P, P, P]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def philox4x32_10(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().oracle_philox4x32_10(_p(c), _p(k), _p(out))
    return out


def uniform(seed, offset, rows, cols, dtype=np.float32):
    out = np.zeros((rows, cols), dtype=dtype)
    fn = lib().oracle_uniform_f32 if dtype == np.float32 else lib().oracle_uniform_f64
    fn(seed & 0xFFFFFFFFFFFFFFFF, offset, rows, cols, _p(out))
    return out


RULES = {"trapezoid": 0, "simpson": 1, "boole": 2, "gauss_legendre": 3}
FAMILIES = {
    "gaussian_peak": 0,
    "oscillatory": 1,
    "product_peak": 2,
    "exp_linear": 3,
    "sin_sum": 4,
    "double_gaussian": 5,
    "constant": 6,
}


def _params(dim, a, u):
    A = np.zeros(max(dim, 1), dtype=np.float64)
    U = np.zeros(max(dim, 1), dtype=np.float64)
    if a is not None:
        A[:dim] = a
    if u is not None:
        U[:dim] = u
    return A, U


def grid_sum(rule, family, dim, n, domain, c0=1.0, c1=0.0, c2=0.0, c3=0.0, a=None, u=None, with_abs=False):
    """FP64 brute-force tensor-product rule with full per-point evaluation.

    with_abs: return (sum w f, sum |w f|) instead of the sum alone."""
    dom = np.ascontiguousarray(np.asarray(domain, dtype=np.float64).reshape(-1))
    A, U = _params(dim, a, u)
    glx = glw = None
    if rule == "gauss_legendre":
        glx, glw = np.polynomial.legendre.leggauss(n)
        glx = np.ascontiguousarray(glx)
        glw = np.ascontiguousarray(glw)
    asum = ctypes.c_double()
    total = lib().oracle_grid_sum(RULES[rule], FAMILIES[family], dim, n, _p(dom), c0, c1, c2, c3, _p(A), _p(U),
                                  _p(glx), _p(glw), ctypes.byref(asum))
    return (total, asum.value) if with_abs else total


def mc_sums(dtype, family, dim, N, seed, offset, domain, c1=0.0, c2=0.0, c3=0.0, a=None, u=None, with_abs=False):
    """(sum f, sum f^2) over the library's sample stream, f in FP64 (no c0);
    with_abs: (sum f, sum f^2, sum |f|)."""
    dom = np.ascontiguousarray(np.asarray(domain, dtype=np.float64).reshape(-1))
    A, U = _params(dim, a, u)
    s1 = ctypes.c_double()
    s2 = ctypes.c_double()
    sa = ctypes.c_double()
    lib().oracle_mc_sum(0 if dtype == "float32" else 1, FAMILIES[family], dim, N, seed & 0xFFFFFFFFFFFFFFFF,
                        offset, _p(dom), 1.0, c1, c2, c3, _p(A), _p(U), ctypes.byref(s1), ctypes.byref(s2),
                        ctypes.byref(sa))
    return (s1.value, s2.value, sa.value) if with_abs else (s1.value, s2.value)
