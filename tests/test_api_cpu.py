"""Host logic and ABI checks that need no GPU."""
import ctypes
import math
import re
import warnings
from pathlib import Path

import numpy as np
import pytest
import torch

import torchquad_amd as tq
from torchquad_amd import _lib
from torchquad_amd.integration.utils import _linspace_with_grads, _setup_integration_domain

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
    # scatter-add: argument checks return before any launch
    fake = ctypes.c_void_p(256)
    assert L.tq_add_at_indices(0, None, 3, None, None, 0, None, None) == 0  # nothing to add
    assert L.tq_add_at_indices(0, fake, 3, None, fake, 4, None, None) == -2
    assert L.tq_add_at_indices(0, fake, 3, fake, fake, -1, None, None) == -1
    assert L.tq_add_at_indices(7, fake, 3, fake, fake, 4, None, None) == -1
    assert b"dtype" in L.tq_last_error()


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


@pytest.mark.parametrize("cls", ["Trapezoid", "Simpson", "Boole", "MonteCarlo"])
def test_calculate_result_error_handling_all_integrators(cls):
    # trapezoid_test.py:177-198, simpson_test.py:177-198, boole_test.py:148-169,
    # monte_carlo_test.py:166-184
    integ = getattr(tq, cls)()
    kw = ({"integration_domain": torch.tensor([[0.0, 1.0]])} if cls == "MonteCarlo" else
          {"dim": 1, "n_per_dim": 101, "hs": torch.tensor([0.01]), "integration_domain": torch.tensor([[0.0, 1.0]])})
    with pytest.raises(ValueError) as exc_info:
        integ.calculate_result(**kw)
    assert "function_values argument not found" in str(exc_info.value)
    assert "Please provide function_values" in str(exc_info.value)
    with pytest.raises(ValueError, match="function_values argument not found"):
        integ.calculate_result()


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("requires_grad", [False, True])
def test_linspace_with_grads(dtype, requires_grad):
    # utils_integration_test.py:47-80
    start = torch.tensor(-2.0, dtype=dtype, requires_grad=requires_grad)
    stop = torch.tensor(-1.0, dtype=dtype, requires_grad=requires_grad)
    g = _linspace_with_grads(start, stop, 10, requires_grad)
    assert g.dtype == dtype and g.shape == (10,)
    assert g[0] == start and g[9] == stop
    assert all(start <= g[i] <= stop for i in range(10))
    assert all(g[i] < g[i + 1] for i in range(9))
    if requires_grad:
        g.sum().backward()
        # d/dstart sum_i (start + i/9 (stop - start)) = 10 - 45/9 = 5, same for stop
        assert abs(float(start.grad) - 5.0) < 1e-6 and abs(float(stop.grad) - 5.0) < 1e-6


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


def test_forms_without_gpu_fail_loudly():
    from torchquad_amd.autograd import gaussian_peak_integrals
    from torchquad_amd.batch import Batch
    from torchquad_amd.serve import IntegrationServer

    if torch.cuda.is_available():
        pytest.skip("checks the no-GPU behaviour")
    with pytest.raises(_lib.NativeUnavailableError):
        IntegrationServer()
    with pytest.raises(_lib.NativeUnavailableError):
        Batch([])
    a = torch.ones(2, 2, dtype=torch.float64)
    dom = torch.tensor([[[0.0, 1.0]] * 2] * 2, dtype=torch.float64)
    with pytest.raises(_lib.NativeUnavailableError):
        gaussian_peak_integrals(a, a, dom, n=4)


def test_batch_plan_host_side():
    """tq_batch_plan is host code: tiles cover every integral exactly once, longest first."""
    from torchquad_amd.batch import n_per_axis

    items = (_lib.BatchItem * 3)()
    spec = [(0, 2, 33), (1, 3, 21), (3, 1, 100000)]
    for it, (rule, dim, n) in zip(items, spec):
        it.rule, it.family, it.dim, it.n = rule, _lib.FAM_GAUSSIAN_PEAK, dim, n
        for k in range(dim):
            it.domain[2 * k], it.domain[2 * k + 1] = 0.0, 1.0
            it.a[k], it.u[k] = 1.0, 0.5
        it.c0 = 1.0
    L = _lib.load_library()
    db, nt = ctypes.c_int64(), ctypes.c_int64()
    assert L.tq_batch_plan_sizes(ctypes.cast(items, ctypes.c_void_p), 3, _lib.TQ_F32, 4096, ctypes.byref(db),
                                 ctypes.byref(nt)) == 0
    import numpy as np

    desc = np.zeros(db.value, dtype=np.uint8)
    tiles = (_lib.BatchTile * nt.value)()
    tb = np.zeros(3, dtype=np.int64)
    tc = np.zeros(3, dtype=np.int64)
    co = np.full(6, -1, dtype=np.int64)  # one spare entry past the 5 the plan writes
    assert L.tq_batch_plan(ctypes.cast(items, ctypes.c_void_p), 3, _lib.TQ_F32, 4096,
                           desc.ctypes.data_as(ctypes.c_void_p), tiles, tb.ctypes.data_as(ctypes.c_void_p),
                           tc.ctypes.data_as(ctypes.c_void_p), co.ctypes.data_as(ctypes.c_void_p)) == 0
    assert tc.sum() == nt.value
    # class 0 holds the one-axis GL integral, class 1 (tables within 4 KB) the
    # other two in LPT order: 21^3 = 9261 points, then 33^2 = 1089
    assert list(np.argsort(tb)) == [2, 1, 0]
    assert list(co[:5]) == [0, tc[2], nt.value, nt.value, nt.value]
    assert co[5] == -1
    for i in range(3):
        ts = [tiles[t] for t in range(tb[i], tb[i] + tc[i])]
        assert all(t.item == i for t in ts)
        assert all(t.lane_wave < t.lane_wave_end for t in ts)
    # every (64-lane group, loop item) pair of an integral is covered exactly once
    for i in range(3):
        seen = set()
        for t in range(tb[i], tb[i] + tc[i]):
            tl = tiles[t]
            for lw in range(tl.lane_wave, tl.lane_wave_end):
                for it in range(tl.t0, tl.t1):
                    assert (lw, it) not in seen
                    seen.add((lw, it))
        n, dim = spec[i][2], spec[i][1]
        assert len(seen) > 0
    assert n_per_axis(_lib.RULE_SIMPSON, 3, 10_000) == 21
    assert n_per_axis(_lib.RULE_BOOLE, 5, 17**5) == 17


def test_serving_helpers():
    from torchquad_amd.serve import Ticket, latency_stats, poisson_arrivals

    offs = poisson_arrivals(2000.0, 20000, seed=0)
    assert np.all(np.diff(offs) > 0)
    assert abs(len(offs) / offs[-1] - 2000.0) < 60.0  # mean rate within 3 %
    ts = []
    for k in range(100):
        t = Ticket("grid", 0, 1)
        t.t_arrival = 0.0
        t._finish(value=1.0)
        t.t_done = (k + 1) * 1e-3
        ts.append(t)
    st = latency_stats(ts)
    assert st["count"] == 100 and st["p50_ms"] == pytest.approx(51.0) and st["p99_ms"] == pytest.approx(100.0)


def _driver(results, sigma2, it, eps_rel=0.0, eps_abs=0.0, sN=1000, inc=1000, fevals=0, N=10**9, max_it=20):
    v = tq.VEGAS()
    v.results, v.sigma2, v._slots = list(results), list(sigma2), []
    v.it, v._eps_rel, v._eps_abs = it, eps_rel, eps_abs
    v._starting_N, v._N_increment, v._nr_of_fevals, v.N, v._max_iterations = sN, inc, fevals, N, max_it
    return v


DRIVER_CASES = [
    ([1.0, 2.0, 4.0], [1.0, 4.0, 16.0], {}),
    ([1.0, 2.0, 6.0], [0.5, 0.0, 2.0], {}),
    ([1.0] * 5, [0.01] * 5, dict(eps_rel=0.1)),
    ([1.0] * 5, [0.01] * 5, dict(eps_abs=0.05)),
    ([1.0] * 5, [0.01] * 5, {}),
    ([1.0] * 5, [5 * 0.0121**2] * 5, dict(eps_rel=0.01, sN=10000, inc=10000)),
    ([0.0] * 5, [0.01] * 5, {}),
    ([3.0, -1.0, 3.0, -1.0, 1.0], [2.0] * 5, {}),
    ([3.0, -1.0, 2.0, 0.0, 1.0], [2.0] * 5, {}),
    ([1.0, 2.0, 1.0, 1.0, 1.0], [1.0, 0.0, 1.0, 1.0, 1.0], {}),
    ([1.0] * 5, [0.01] * 5, dict(fevals=5001, N=15000)),
    ([1.0] * 5, [0.01] * 5, dict(max_it=19)),
]


@pytest.mark.parametrize("case", range(len(DRIVER_CASES)))
def test_vegas_driver_host_logic_matches_oracle(case):
    """EQ 30-32 and the abort / N schedule of the VEGAS driver (vegas.py:161-209,
    318-362) against the oracle's restatement, whose cases are pinned by hand."""
    from oracle import vegas as ov

    r, s, kw = DRIVER_CASES[case]
    v = _driver(r, s, 5, **kw)
    assert float(v._get_result()) == ov._result(r, s)
    assert float(v._get_error()) == ov._error(s)
    assert float(v._get_chisq()) == ov._chisq(r, s)
    args = dict(eps_rel=0.0, eps_abs=0.0, sN=1000, inc=1000, fevals=0, N=10**9, max_it=20)
    args.update(kw)
    stop, sN = ov.abort_check(5, r, s, args["eps_rel"], args["eps_abs"], args["sN"], args["inc"], args["fevals"],
                              args["N"], args["max_it"])
    assert v._check_abort_conditions() == stop
    assert v._starting_N == sN
    if not stop:
        assert v.results == [] and v.sigma2 == []
    # not a fifth iteration: never stops, nothing changes
    v = _driver(r, s, 4, **kw)
    assert v._check_abort_conditions() is False and v._starting_N == args["sN"]


def test_noalloc_rejects_grids_beyond_its_scratch():
    """tq_grid_integrate_noalloc refuses, before any device call, a grid whose
    block partials exceed its 1 MiB static scratch."""
    L = _lib.load_library()
    g = _lib.GridDesc()
    g.dtype, g.rule, g.dim, g.n, g.items_per_wave = 0, 0, 6, 1001, 1
    need = L.tq_grid_workspace_size(ctypes.byref(g))
    assert need > (1 << 20)
    f = _lib.Integrand()
    f.family, f.dim, f.c0 = 0, 6, 1.0
    fake = ctypes.c_void_p(256)
    rc = L.tq_grid_integrate_noalloc(ctypes.byref(g), ctypes.byref(f), fake, fake, None, None)
    assert rc == -3
    assert b"scratch" in L.tq_last_error()
    g.n, g.dim, g.items_per_wave = 9, 3, 0
    assert L.tq_grid_workspace_size(ctypes.byref(g)) <= (1 << 20)
    g.rule = 2  # Boole needs n = 1 + 4k
    g.n = 10
    assert L.tq_grid_integrate_noalloc(ctypes.byref(g), ctypes.byref(f), fake, fake, None, None) == -1


def test_batch_run_rejects_bad_gl_tables_on_the_host():
    """tq_batch_run checks the Gauss-Legendre table arguments before any
    device call: a negative gl_len is TQ_E_ARG, a positive one with NULL
    node/weight tables TQ_E_NULL."""
    L = _lib.load_library()
    fake = ctypes.c_void_p(256)
    offs = np.zeros(5, dtype=np.int64)

    def run(gx, gw, gl_len):
        return L.tq_batch_run(0, fake, 1, fake, 1, fake, fake, gx, gw, None, gl_len, fake, fake, fake,
                              offs.ctypes.data_as(ctypes.c_void_p), None)

    assert run(fake, fake, -1) == -1
    assert b"gl_len" in L.tq_last_error()
    assert run(None, fake, 10) == -2
    assert run(fake, None, 10) == -2
    assert b"gl_len > 0" in L.tq_last_error()


def test_gaussian_rule_classification():
    """Which Gaussian rules run on the kernels: hooks classified on probes
    (gaussian.py:13-25, tutorial.rst:1198-1236)."""
    import scipy.special

    from torchquad_amd.integration import gaussian as gm

    assert tq.GaussLegendre()._tq_mode() == gm.AFFINE
    assert tq.GaussHermite()._tq_mode() == gm.UNIT
    assert tq.Gaussian()._tq_mode() is None  # identity map with the (b-a)/2 scale

    class UserHermite(tq.Gaussian):  # the tutorial's subclass, verbatim hooks
        def __init__(self):
            super().__init__()
            self._root_fn = scipy.special.roots_hermite

        @staticmethod
        def _apply_composite_rule(cur_dim_areas, dim, hs, domain):
            for _ in range(dim):
                cur_dim_areas = torch.sum(cur_dim_areas, dim=len(cur_dim_areas.shape) - 1)
            return cur_dim_areas

    class Lobatto(tq.GaussLegendre):  # another root function, Gauss-Legendre hooks
        def __init__(self):
            super().__init__()
            self._root_fn = lambda n: np.polynomial.legendre.leggauss(n)

    class Doubled(tq.GaussLegendre):  # a composite rule the kernels do not implement
        @staticmethod
        def _apply_composite_rule(cur_dim_areas, dim, hs, domain):
            return 2.0 * torch.sum(cur_dim_areas.reshape(*cur_dim_areas.shape[:-dim], -1), dim=-1)

    class OwnRoots(tq.GaussLegendre):
        def _roots(self, N, backend, requires_grad=False):
            return super()._roots(N, backend, requires_grad)

    class Broken(tq.GaussLegendre):  # a hook that fails on the probes runs its own code
        def _resize_roots(self, integration_domain, roots):
            raise RuntimeError("only on the GPU")

    assert UserHermite()._tq_mode() == gm.UNIT
    assert Lobatto()._tq_mode() == gm.AFFINE
    assert Doubled()._tq_mode() is None
    assert OwnRoots()._tq_mode() is None
    assert Broken()._tq_mode() is None
    unit = UserHermite()._tq_domain(torch.tensor([[0.0, 3.0], [2.0, 5.0]], dtype=torch.float64))
    assert unit.tolist() == [[-1.0, 1.0], [-1.0, 1.0]] and unit.dtype == torch.float64


def test_param_batch_integrands_host_side():
    """Parameter batches: shapes, descriptors and the [N, *S] torch definition."""
    a = np.arange(1.0, 7.0).reshape(3, 2)
    f = tq.GaussianPeak(a, [0.1, 0.2])
    assert f.batch_shape == (3,) and f.batch_size == 3 and f.dim == 2
    ds = f.descriptors()
    assert [d.a[1] for d in ds] == [2.0, 4.0, 6.0] and all(d.u[0] == 0.1 for d in ds)
    with pytest.raises(ValueError, match="parameter batch"):
        f.descriptor()
    x = torch.rand(11, 2, dtype=torch.float64)
    y = f(x)
    assert y.shape == (11, 3)
    for s in range(3):
        ref = tq.GaussianPeak(a[s], [0.1, 0.2])(x)
        assert torch.allclose(y[:, s], ref, rtol=1e-15, atol=0)
    g = tq.Oscillatory(np.ones((2, 3, 1)), w=np.array([[0.1, 0.2, 0.3], [0.4, 0.5, 0.6]]), dim=1)
    assert g.batch_shape == (2, 3) and g(torch.rand(5, 1)).shape == (5, 2, 3)
    assert [d.c1 for d in g.descriptors()] == [0.1, 0.2, 0.3, 0.4, 0.5, 0.6]
    h = tq.DoubleGaussian(3, width=[0.1, 0.2])
    ds = h.descriptors()
    assert ds[1].c3 == pytest.approx(25.0) and ds[0].c0 == pytest.approx(0.5 * (0.1 * math.sqrt(math.pi)) ** -3)
    assert tq.GaussianPeak([1.0, 2.0], [0.0, 0.0]).batch_shape == ()
    assert f.key() != tq.GaussianPeak(a + 1e-9, [0.1, 0.2]).key()
    with pytest.raises(ValueError, match="broadcast"):
        tq.GaussianPeak(np.ones((3, 2)), np.zeros((4, 2)))


def test_serving_item_records_match_the_c_layout():
    """numpy records of submit_many have tq_batch_item's layout; vectorised N
    rounding equals the per-integral one; the Genz arrays carry the same
    descriptors as the Genz items."""
    from torchquad_amd.batch import n_per_axis
    from torchquad_amd.serve import ITEM_DTYPE, n_per_axis_array
    from torchquad_amd.workloads import genz_arrays, genz_batch

    assert ITEM_DTYPE.itemsize == ctypes.sizeof(_lib.BatchItem)
    for name, _ in _lib.BatchItem._fields_:
        assert ITEM_DTYPE.fields[name][1] == getattr(_lib.BatchItem, name).offset, name
    rng = np.random.default_rng(0)
    rule = rng.integers(0, 4, 500)
    dim = rng.integers(1, 7, 500)
    N = np.maximum((10 ** rng.uniform(1.5, 6.0, 500)).astype(np.int64), 2**dim)
    vec = n_per_axis_array(rule, dim, N)
    assert vec.tolist() == [n_per_axis(int(r), int(d), int(m)) for r, d, m in zip(rule, dim, N)]
    items = genz_batch(50, seed=3)
    arr = genz_arrays(50, seed=3)
    for i, it in enumerate(items):
        d = it["fn"].descriptor(it["dim"])
        assert arr["family"][i] == d.family and arr["c"][i].tolist() == [d.c0, d.c1, d.c2, d.c3]
        assert arr["a"][i].tolist() == list(d.a) and arr["u"][i].tolist() == list(d.u)


class _HostOnlyServer:
    """The submission half of IntegrationServer without a device: requests
    are packed and kept, nothing is launched."""

    def __new__(cls):
        from torchquad_amd.serve import IntegrationServer

        class _S(IntegrationServer):
            def __init__(self):
                self.dt, self.tile_points, self.hbm_budget, self.queued = _lib.TQ_F32, 4096, 1 << 40, []

            def _enqueue(self, r):
                self.queued.append(r)

        return _S()


def test_submit_packing_on_the_host():
    """submit_grid's record equals the field-by-field tq_batch_item; submit_many
    takes its count from any per-integral argument (here only n) and packs
    the same records as one submit_grid per integral."""
    from torchquad_amd.serve import ITEM_DTYPE

    s = _HostOnlyServer()
    f = tq.GaussianPeak([3.0, 2.0, 1.0], [0.4, 0.5, 0.6])
    dom = [[0.0, 1.0], [0.1, 0.9], [0.2, 0.7]]
    s.submit_grid("simpson", f, 3, N=100000, domain=dom)
    ref = np.zeros(1, ITEM_DTYPE)
    d = f.descriptor(3)
    ref["rule"], ref["family"], ref["dim"], ref["n"] = _lib.RULE_SIMPSON, d.family, 3, 45
    ref["domain"][0, :6] = np.asarray(dom).reshape(-1)
    ref["c0"], ref["c1"], ref["c2"], ref["c3"] = d.c0, d.c1, d.c2, d.c3
    ref["a"][0] = np.ctypeslib.as_array(d.a)
    ref["u"][0] = np.ctypeslib.as_array(d.u)
    assert s.queued[0].items.tobytes() == ref.tobytes()
    assert s.queued[0].gl_n.tolist() == [0]

    ns = [33, 2000, 33, 2500]
    g = tq.GaussianPeak([3.0], [0.4])
    s.submit_many(_lib.RULE_GAUSS_LEGENDRE, _lib.FAM_GAUSSIAN_PEAK, 1, [[0.0, 1.0]], [3.0], [0.4], n=ns)
    many = s.queued[-1]
    assert len(many.items) == 4 and many.gl_n.tolist() == ns and many.gl_unique.tolist() == [33, 2000, 2500]
    for n in ns:
        s.submit_grid("gauss_legendre", g, 1, n=n, domain=[[0.0, 1.0]])
    singles = np.concatenate([r.items for r in s.queued[-4:]])
    assert many.items.tobytes() == singles.tobytes()


def test_probe_rates_rejects_bad_arguments_on_the_host():
    L = _lib.load_library()
    fake = ctypes.c_void_p(256)
    assert L.tq_probe_rates(4, 8, 1, fake, None) == -1
    assert L.tq_probe_rates(0, 0, 1, fake, None) == -1
    assert L.tq_probe_rates(0, 8, 0, fake, None) == -1
    assert L.tq_probe_rates(0, 8, 1, None, None) == -2
