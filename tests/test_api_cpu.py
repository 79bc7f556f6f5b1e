"""Host logic and ABI checks that need no GPU."""
import ctypes
import re
import warnings
from pathlib import Path

import pytest
import torch

import torchquad_amd as tq
from torchquad_amd import _lib
from torchquad_amd.integration.utils import _setup_integration_domain

HEADERS = sorted((Path(__file__).resolve().parent.parent / "include").glob("*.h"))


def declared_symbols():
    names = set()
    for h in HEADERS:
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(tq_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_library_loads_and_exports_every_declared_symbol():
    L = _lib.load_library()
    syms = declared_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(L, s), f"{s} declared in include/ but not exported"
        assert s in _lib.SIGNATURES, f"{s} has no ctypes signature"
    assert L.tq_version() >= 100


def test_invalid_arguments_return_error_codes():
    L = _lib.load_library()
    g = _lib.GridDesc()
    g.dtype, g.rule, g.dim, g.n = 0, 2, 3, 6  # Boole needs n = 1 + 4k
    f = _lib.Integrand()
    f.family, f.dim = 0, 3
    rc = L.tq_grid_integrate(ctypes.byref(g), ctypes.byref(f), None, None, 0, None, None, None)
    assert rc < 0
    g.n = 5
    rc = L.tq_grid_integrate(ctypes.byref(g), ctypes.byref(f), None, None, 0, None, None, None)
    assert rc == -2  # NULL domain
    assert b"NULL" in L.tq_last_error()
    g.rule = 3  # Gauss-Legendre without tables
    rc = L.tq_grid_integrate(ctypes.byref(g), ctypes.byref(f), ctypes.c_void_p(1), None, 0, ctypes.c_void_p(1), None, None)
    assert rc == -2
    m = _lib.McDesc()
    m.dtype, m.dim, m.N = 0, 0, 10
    assert L.tq_mc_workspace_size(ctypes.byref(m)) == 0
    assert L.tq_uniform(1, 0, 5, 1, 0, None, None) < 0


def test_workspace_sizes_positive():
    L = _lib.load_library()
    g = _lib.GridDesc()
    g.dtype, g.rule, g.dim, g.n = 0, 2, 5, 17
    assert L.tq_grid_workspace_size(ctypes.byref(g)) > 0
    assert L.tq_grid_reduce_workspace_size(ctypes.byref(g), 1) > 0
    m = _lib.McDesc()
    m.dtype, m.dim, m.N = 0, 8, 10**8
    assert L.tq_mc_workspace_size(ctypes.byref(m)) > 0


def test_adjust_N_warnings_match_reference():
    with pytest.warns(UserWarning, match="cannot be even"):
        assert tq.Simpson._adjust_N(3, 1076890) == 101**3
    with pytest.warns(UserWarning, match="N = 1 \\+ 4n"):
        assert tq.Boole._adjust_N(3, 1076890) == 101**3
    with pytest.warns(UserWarning, match="lower than 3"):
        assert tq.Simpson._adjust_N(2, 4) == 9
    with pytest.warns(UserWarning, match="lower than 5"):
        assert tq.Boole._adjust_N(1, 3) == 5
    assert tq.Simpson._get_minimal_N(4) == 81 and tq.Boole._get_minimal_N(2) == 25


def test_domain_setup_and_errors():
    d = _setup_integration_domain(2, None, None)
    assert tuple(d.shape) == (2, 2) and d.dtype == torch.get_default_dtype()
    with pytest.raises(ValueError, match="unexpected shape"):
        _setup_integration_domain(3, [[0, 1], [0, 1]], None)
    with pytest.raises(ValueError, match="Unsupported numerical backend"):
        _setup_integration_domain(1, [[0, 1]], "jax")
    with pytest.raises(ValueError):
        tq.Trapezoid()._check_inputs(dim=0)
    with pytest.raises(ValueError, match="positive integer"):
        tq.Trapezoid()._check_inputs(N=1.5)
    with pytest.raises(ValueError, match="must match"):
        tq.Trapezoid()._check_inputs(dim=2, integration_domain=[[0, 1]])
    with pytest.raises(ValueError, match="invalid boundary values"):
        tq.Trapezoid()._check_inputs(dim=1, integration_domain=torch.tensor([[1.0, 0.0]]))
    with pytest.raises(ValueError, match="invalid shape"):
        tq.Trapezoid()._check_inputs(integration_domain=torch.zeros(1, 3, 2))


def test_calculate_result_missing_values_message():
    with pytest.raises(ValueError, match="function_values argument not found"):
        tq.Trapezoid().calculate_result(dim=1, n_per_dim=101, hs=torch.tensor([0.01]),
                                        integration_domain=torch.tensor([[0.0, 1.0]]))
    with pytest.raises(ValueError, match="Please provide function_values"):
        tq.Trapezoid().calculate_result()


def test_grid_checks():
    with pytest.raises(ValueError, match="N has to be > 1"):
        tq.Trapezoid()._grid_n(1, 1)


def test_seed_and_rng_exclusive():
    with pytest.raises(ValueError, match="cannot both"):
        tq.MonteCarlo().calculate_sample_points(10, torch.tensor([[0.0, 1.0]]), seed=1, rng=tq.RNG(seed=2))


@pytest.mark.skipif(torch.cuda.is_available(), reason="checks the no-GPU error path")
def test_compute_without_gpu_fails_loudly():
    with pytest.raises(_lib.NativeUnavailableError):
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            tq.Boole().integrate(tq.GaussianPeak([1, 2], [0, 0]), 2, N=25, integration_domain=[[0, 1], [0, 1]])
    with pytest.raises(_lib.NativeUnavailableError):
        tq.MonteCarlo().integrate(lambda x: x[:, 0], 1, N=10, integration_domain=[[0, 1]], seed=0)


def test_builtin_torch_definitions_match_oracle():
    import numpy as np
    from oracle import families as fam

    rng = np.random.default_rng(0)
    x = rng.uniform(-1, 1, (50, 3))
    a, u, c = rng.uniform(1, 3, 3), rng.uniform(-0.5, 0.5, 3), rng.uniform(0, 2, 3)
    xt = torch.tensor(x)
    pairs = [
        (tq.GaussianPeak(a, u), fam.gaussian_peak(x, a, u)),
        (tq.Oscillatory(c, 0.3), fam.oscillatory(x, c, 0.3)),
        (tq.ProductPeak(a, u), fam.product_peak(x, a, u)),
        (tq.ExpLinear(3, a, 0.1), fam.exp_linear(x, a, 0.1)),
        (tq.SinSum(3, a, u), fam.sin_sum(x, a, u)),
        (tq.DoubleGaussian(3, 0.2), fam.double_gaussian(x, 0.2)),
        (tq.Constant(3, 2.5), np.full(50, 2.5)),
    ]
    for f, ref in pairs:
        np.testing.assert_allclose(f(xt).numpy(), ref, rtol=1e-12, atol=1e-14)
