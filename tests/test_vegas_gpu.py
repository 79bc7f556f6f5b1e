"""VEGAS on the GPU vs the oracle restatement and the reference's checks."""
import math
import warnings
from unittest.mock import patch

import numpy as np
import pytest
import torch

import torchquad_amd as tq
from oracle import families as fam
from oracle import native
from oracle import vegas as ov

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _quiet():
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        yield


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_map_golden_vectors(ref_vectors, dtype, dev):
    g = ref_vectors["vegas_map_interval"]
    m = tq.VEGASMap(g["N_intervals"], 3, "torch", dtype)
    y = torch.tensor(g["y"], dtype=dtype, device=dev)
    assert torch.equal(m._get_interval_ID(y).cpu(), torch.tensor(g["ID"]))
    assert float(torch.max(torch.abs(m._get_interval_offset(y).cpu() - torch.tensor(g["offset"], dtype=dtype)))) <= 6e-5
    assert float(torch.max(torch.abs(m.get_X(y) - y))) <= 3e-7
    gc = ref_vectors["vegas_map_counts"]
    y1 = torch.linspace(0.0, gc["y_max"], gc["N_per_dim"], dtype=dtype, device=dev)
    yy = torch.stack([a.ravel() for a in torch.meshgrid(y1, y1, y1, indexing="ij")], 1)
    f = torch.prod(torch.exp(yy), dim=1)
    jac = m.get_Jac(yy)
    assert jac.dtype == dtype and float(torch.max(torch.abs(jac - 1.0))) < 1e-14
    m.accumulate_weight(yy, (f * jac) ** 2)
    w = m.weights
    assert float(torch.min(w[:, 1:] - w[:, :-1])) > 0.0
    assert int(torch.max(torch.abs(m.counts - gc["count"]))) == 0
    gs = ref_vectors["vegas_smooth_map"]
    sw = tq.VEGASMap._smooth_map(torch.tensor(gs["weights"], dtype=dtype, device=dev),
                                 torch.ones((2, 6), dtype=torch.int64, device=dev), 0.5)
    assert float(torch.max(torch.abs(sw.cpu() - torch.tensor(gs["expected"], dtype=dtype)))) <= 3e-7
    m.update_map()
    xe, dxe = m.x_edges, m.dx_edges
    assert float(torch.max(torch.abs(xe[:, [0, -1]] - torch.tensor([[0.0, 1.0]] * 3, dtype=dtype, device=dev)))) == 0
    assert float(torch.max(torch.abs(torch.sum(dxe, 1) - 1.0))) < 3e-7
    assert float(torch.min(dxe)) > 0 and float(torch.max(dxe[:, 1:] - dxe[:, :-1])) < 0
    assert float(torch.max(torch.abs(m.get_X(yy)[0]))) == 0.0


def test_map_update_matches_oracle(dev):
    rng = np.random.default_rng(3)
    for N, dim in ((2, 1), (7, 3), (50, 6)):
        m = tq.VEGASMap(N, dim, "torch", torch.float64)
        om = ov.Map(N, dim)
        w = rng.uniform(0.0, 3.0, (dim, N))
        c = rng.integers(0, 5, (dim, N))
        w[c == 0] = 0.0
        m._dev.weights.copy_(torch.tensor(w))
        m._dev.counts.copy_(torch.tensor(c, dtype=torch.float64))
        om.weights, om.counts = w.copy(), c.copy()
        sw_dev = tq.VEGASMap._smooth_map(torch.tensor(w, device=dev), torch.tensor(c, device=dev), 0.5)
        sw_ref = ov.Map.smooth(w, c, 0.5)
        if sw_ref is None:
            assert sw_dev is None
            continue
        np.testing.assert_allclose(sw_dev.cpu().numpy(), sw_ref, rtol=1e-14, atol=1e-15)
        m.update_map()
        om.update_map()
        np.testing.assert_allclose(m.x_edges.cpu().numpy(), om.x_edges, rtol=0, atol=1e-14)


def test_zero_weights_cannot_update(dev):
    m = tq.VEGASMap(4, 2, "torch", torch.float64)
    before = m.x_edges.clone()
    with pytest.warns(RuntimeWarning, match="Cannot update the VEGASMap"):
        m.update_map()
    assert torch.equal(before, m.x_edges)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_stratification_bit_exact(dtype, dev, ref_vectors):
    g = ref_vectors["vegas_strat_fresh"]
    rng = tq.RNG(seed=11)
    st = tq.VEGASStratification(g["N_increment"], g["dim"], rng, "torch", dtype)
    ost = ov.Strat(g["N_increment"], g["dim"], np.float32 if dtype == torch.float32 else np.float64)
    assert st.N_strat == ost.N_strat == g["N_strat"]
    nh = st.get_NH(g["nevals_exp"])
    assert nh.dtype == torch.int64 and torch.all(nh == g["nh"])
    y = st.get_Y(nh)
    u = native.uniform(11, 0, int(nh.sum()), 3, np.float32 if dtype == torch.float32 else np.float64)
    assert np.array_equal(y.cpu().numpy(), ost.get_Y(nh.cpu().numpy(), u))
    f = torch.prod(torch.exp(y), dim=1)
    jf, jf2 = st.accumulate_weight(nh, f)
    ojf, ojf2 = ost.accumulate_weight(nh.cpu().numpy(), f.double().cpu().numpy())
    np.testing.assert_allclose(jf.double().cpu().numpy(), ojf, rtol=1e-12 if dtype == torch.float64 else 1e-6)
    st.update_DH()
    ost.update_DH()
    np.testing.assert_allclose(st._dev.dh.cpu().numpy(), ost.dh, rtol=1e-12)
    # given the same dh, get_NH is bit-exact for several expected counts
    ost.dh = st._dev.dh.cpu().numpy()
    for n_exp in (4000, 123457, 10**6):
        assert np.array_equal(st.get_NH(n_exp).cpu().numpy(), ost.get_NH(n_exp))
    assert int(st.get_NH(4000)[-1]) > int(st.get_NH(4000)[0])


# (dim, bins): the sample kernel's instantiations -- dimension bound 8 or 16,
# wave histograms in LDS (dim * bins <= 768) or global atomics (default bins)
@pytest.mark.parametrize("n_strat", [724, 725])
def test_nh_offsets_both_scan_forms(n_strat, dev):
    """get_NH and the sample offsets around the switch between the two scan
    forms: 724^2 cubes give 256 nh blocks of 2048 cubes (block sums added by
    the offsets kernel itself), 725^2 give 257 (separate scan kernel).  nh
    bit-exact against the oracle for a non-uniform dh, offsets exactly the
    exclusive prefix sums of nh."""
    N = 4 * n_strat * n_strat
    st = tq.VEGASStratification(N, 2, tq.RNG(seed=3), "torch", torch.float64)
    ost = ov.Strat(N, 2)
    assert st.N_strat == ost.N_strat == n_strat
    dh = np.random.default_rng(n_strat).exponential(1.0, n_strat * n_strat)
    dh /= dh.sum()
    st._dev.dh.copy_(torch.tensor(dh))
    ost.dh = dh.copy()
    for n_exp in (10**6, 3 * 10**7):
        nh = st.get_NH(n_exp).cpu().numpy()
        assert np.array_equal(nh, ost.get_NH(n_exp))
        off = st._dev.offsets.cpu().numpy()
        assert off[0] == 0 and np.array_equal(off[1:], np.cumsum(nh))


@pytest.mark.parametrize("dim,bins", [(3, None), (9, None), (3, 40), (9, 40)])
def test_fused_end_to_end_fp64_matches_oracle(dim, bins, dev):
    N, seed = 60000, 5
    dom = ([[0.0, 1.0], [-0.5, 1.0], [0.2, 0.9]] * 3)[:dim]
    f = tq.DoubleGaussian(dim, width=0.3)
    v = tq.VEGAS()
    val = float(v.integrate(f, dim, N=N, integration_domain=torch.tensor(dom, dtype=torch.float64, device=dev),
                            seed=seed, N_intervals=bins))
    r = ov.run(lambda x: fam.double_gaussian(x, 0.3), dim, N, dom, seed=seed, N_intervals=bins)
    assert v.map.N_intervals == r["map"].N
    assert v._nr_of_fevals == r["fevals"]
    assert abs(val - r["result"]) <= 1e-10 * abs(r["result"])
    # map state after the run
    np.testing.assert_allclose(v._dev.x_edges.cpu().numpy(), r["map"].x_edges, atol=1e-12)


def test_user_callable_end_to_end_fp64_matches_oracle(dev):
    dim, N, seed = 2, 20000, 7
    dom = [[0.0, 2.0], [-1.0, 1.0]]
    fn = lambda x: torch.exp(-((x[:, 0] - 1.2) ** 2) * 4.0 - (x[:, 1] ** 2) * 2.0)  # noqa: E731
    v = tq.VEGAS()
    val = float(v.integrate(fn, dim, N=N, integration_domain=torch.tensor(dom, dtype=torch.float64, device=dev),
                            seed=seed))
    r = ov.run(lambda x: np.exp(-((x[:, 0] - 1.2) ** 2) * 4.0 - (x[:, 1] ** 2) * 2.0), dim, N, dom, seed=seed)
    assert v._nr_of_fevals == r["fevals"]
    assert abs(val - r["result"]) <= 1e-10 * abs(r["result"])
    # fused built-in of the same function agrees with the callable path
    g = tq.GaussianPeak([2.0, math.sqrt(2.0)], [1.2, 0.0])
    val2 = float(tq.VEGAS().integrate(g, dim, N=N, integration_domain=torch.tensor(dom, dtype=torch.float64,
                                                                                    device=dev), seed=seed))
    assert abs(val2 - val) <= 1e-9 * abs(val)


def test_fp32_statistics(configs, dev):
    c = configs["C3"]
    v = tq.VEGAS()
    f = tq.DoubleGaussian(6, width=0.1)
    dom = torch.tensor(c["domain"], dtype=torch.float32, device=dev)
    val = float(v.integrate(f, 6, N=2_000_000, integration_domain=dom, seed=0, N_intervals=50))
    err = v._get_error()
    assert abs(val - c["analytic"]) <= max(5 * err, 0.01 * c["analytic"]), (val, c["analytic"], err)


def test_reference_special_cases(ref_vectors, dev):
    g = ref_vectors["vegas_special"]
    v = tq.VEGAS()
    val = v.integrate(lambda x: x[:, 0] * 0.0, 2, N=10000, integration_domain=[[0.0, 3.0]] * 2, seed=0)
    assert float(val) == 0.0
    d64 = torch.tensor(g["constant_domain"], dtype=torch.float64, device=dev)
    val = v.integrate(lambda x: x[:, 0] * 0.0 + 10.0, 2, N=10000, integration_domain=d64, seed=0)
    assert abs(float(val) - 90.0) < 1e-13
    val = v.integrate(tq.Constant(2, 10.0), 2, N=10000, integration_domain=d64, seed=0)
    assert abs(float(val) - 90.0) < 1e-13


class ModifiedRNG(tq.RNG):
    """Replaces some uniforms by 0.0 and 1.0 (vegas_test.py:143-156)."""

    def __init__(self, *args, **kargs):
        super().__init__(*args, **kargs)
        rng_uniform = self.uniform
        self.uniform = lambda *a, **k: self.modify_numbers(rng_uniform(*a, **k))

    def modify_numbers(self, numbers):
        zeros = torch.zeros_like(numbers)
        ones = torch.ones_like(numbers)
        return torch.where(numbers < 0.5, numbers * 2.0, torch.where(numbers < 0.75, zeros, ones))


def test_edge_random_numbers(dev):
    with patch("torchquad_amd.integration.vegas.RNG", ModifiedRNG):
        v = tq.VEGAS()
        val = v.integrate(lambda x: torch.sum(x, dim=1), 2, N=10000, integration_domain=[[0.0, 1.0]] * 2, seed=0)
    assert isinstance(v.rng, ModifiedRNG)
    assert abs(float(val) - 1.0) < 0.1


def test_reference_accuracy_checks(dev):
    # vegas_test.py:71-140
    for dtype in (torch.float32, torch.float64):
        dom = torch.tensor([[1.0, 5.0], [-4.0, 4.0], [2.0, 6.0]], dtype=dtype, device=dev)

        def cube_peak(x):
            return torch.prod((x >= 3.0) * (x < 4.0), dim=1).to(dtype) + 0.001

        ref = float(torch.prod(dom[:, 1] - dom[:, 0])) * 0.001 + 1.0
        for seed in [0, 1, 2, 3, 41317]:
            val = tq.VEGAS().integrate(cube_peak, 3, N=30000, integration_domain=dom, seed=seed)
            assert abs(float(val) - ref) < 0.03
        pd = 100.0
        dom2 = torch.tensor([[1.0, 1.0 + pd], [-4.0, -4.0 + pd]], dtype=dtype, device=dev)

        def diag(x):
            return torch.exp(torch.sum(dom2[:, 0] - x, dim=1)) + torch.exp(torch.sum(x - dom2[:, 1], dim=1))

        ref2 = 2.0 - 4.0 * math.exp(-pd) + 2.0 * math.exp(-2 * pd)
        for seed in [0, 1, 2, 3, 41317]:
            val = tq.VEGAS().integrate(diag, 2, N=30000, integration_domain=dom2, seed=seed)
            assert abs(float(val) - ref2) < 0.03


def test_vegas_gradients(ref_vectors, dev):
    G = ref_vectors["gradients"]
    v = tq.VEGAS()
    dom = torch.tensor(G["v_function_domain"]["domain"], dtype=torch.float64, device=dev, requires_grad=True)
    val = v.integrate(lambda x: 2 * torch.abs(x), 1, N=99997, integration_domain=dom, seed=0)
    val.backward()
    assert abs(float(val) - 2.0) < 1e-2
    np.testing.assert_array_less(np.abs(dom.grad.cpu().numpy() - np.array([[-2.0, 2.0]])), 2e-2)
    dom = torch.tensor(G["poly_domain"]["domain"], dtype=torch.float64, device=dev, requires_grad=True)
    val = v.integrate(lambda x: 1.0 + 2.0 * x[:, 0] + 3.0 * x[:, 1] ** 2, 2, N=99997, integration_domain=dom, seed=0)
    val.backward()
    assert abs(float(val) - 12.0) < 8e-2
    np.testing.assert_array_less(np.abs(dom.grad.cpu().numpy() - np.array(G["poly_domain"]["grad"])), 0.1)
    param = torch.tensor([1.0, 2.0, 3.0], dtype=torch.float64, device=dev, requires_grad=True)
    val = v.integrate(lambda x: param[0] + param[1] * x[:, 0] + param[2] * x[:, 1] ** 2, 2, N=99997,
                      integration_domain=[[0.0, 1.0], [0.0, 2.0]], seed=0)
    val.backward()
    np.testing.assert_array_less(np.abs(param.grad.cpu().numpy() - np.array([2.0, 1.0, 8.0 / 3.0])), 5e-2)


def test_warm_start_and_state_dict(dev):
    """state_dict() -> integrate(warm_start=...): the map and dh are carried over exactly."""
    f = tq.DoubleGaussian(3, width=0.2)
    dom = torch.tensor([[0.0, 1.0]] * 3, dtype=torch.float64, device=dev)
    analytic = fam.double_gaussian_integral(3, [0.0] * 3, [1.0] * 3, width=0.2)
    v1 = tq.VEGAS()
    v1.integrate(f, 3, N=60_000, integration_domain=dom, seed=1, N_intervals=20)
    st = v1.state_dict()
    v2 = tq.VEGAS()
    v2._setup(f, 3, 60_000, dom, 2, None, True, 0, 0, 20, None, 20, None, st)
    assert torch.equal(v2._dev.x_edges, st["x_edges"]) and torch.equal(v2._dev.dh, st["dh"])
    v3 = tq.VEGAS()
    r3 = float(v3.integrate(f, 3, N=60_000, integration_domain=dom, seed=2, N_intervals=20, warm_start=st,
                            use_warmup=False))
    assert abs(r3 - analytic) <= max(5 * v3._get_error(), 1e-2 * analytic)
    with pytest.raises(ValueError):
        tq.VEGAS().integrate(f, 3, N=60_000, integration_domain=dom, seed=2, N_intervals=10, warm_start=st)


def test_reused_state_matches_fresh(dev):
    """A VEGAS object reuses its device state across integrate() calls of the same shape (one fill
    instead of new buffers): every call gives exactly what a fresh object gives, also after a larger
    call grew the sample buffers and for a user callable."""
    f = tq.DoubleGaussian(3, width=0.2)
    dom = torch.tensor([[0.0, 1.0]] * 3, dtype=torch.float64, device=dev)

    def user(x):
        return torch.exp(-torch.sum((x - 0.5) ** 2, dim=1) / 0.08)

    v = tq.VEGAS()
    for fn, N, seed in ((f, 60_000, 1), (f, 200_000, 2), (f, 60_000, 1), (user, 60_000, 3), (f, 60_000, 1)):
        again = float(v.integrate(fn, 3, N=N, integration_domain=dom, seed=seed, N_intervals=20))
        fresh = float(tq.VEGAS().integrate(fn, 3, N=N, integration_domain=dom, seed=seed, N_intervals=20))
        assert again == fresh, (N, seed, again, fresh)


def test_n_strat_override(dev):
    f = tq.DoubleGaussian(3, width=0.2)
    dom = torch.tensor([[0.0, 1.0]] * 3, dtype=torch.float64, device=dev)
    v = tq.VEGAS()
    r = float(v.integrate(f, 3, N=60_000, integration_domain=dom, seed=4, N_intervals=20, N_strat=4))
    assert v.strat.N_strat == 4 and v._dev.n_cubes == 64
    analytic = fam.double_gaussian_integral(3, [0.0] * 3, [1.0] * 3, width=0.2)
    assert abs(r - analytic) <= max(5 * v._get_error(), 1e-2 * analytic)


def test_served_vegas_warm_start(dev):
    from torchquad_amd.serve import IntegrationServer

    f = tq.DoubleGaussian(3, width=0.2)
    server = IntegrationServer(dev, dtype=torch.float64)
    t1 = server.submit_vegas(f, 3, N=60_000, domain=[[0.0, 1.0]] * 3, seed=1, N_intervals=20)
    server.run_until_idle(timeout=60)
    t2 = server.submit_vegas(f, 3, N=60_000, domain=[[0.0, 1.0]] * 3, seed=2, N_intervals=20, warm_start=t1.state,
                             use_warmup=False)
    server.run_until_idle(timeout=60)
    dom = torch.tensor([[0.0, 1.0]] * 3, dtype=torch.float64, device=dev)
    ref = float(tq.VEGAS().integrate(f, 3, N=60_000, integration_domain=dom, seed=2, N_intervals=20,
                                     warm_start=t1.state, use_warmup=False))
    assert t2.result() == ref


def test_add_at_indices_reference_cases(ref_vectors, dev):
    # utils_integration_test.py:83-115, through the HIP kernel
    from torchquad_amd.integration.utils import _add_at_indices

    for dtype in (torch.float32, torch.float64):
        target = torch.zeros(500, dtype=dtype, device=dev)
        indices = torch.arange(500, device=dev)
        source = torch.ones(500, dtype=dtype, device=dev)
        _add_at_indices(target, indices, source, is_sorted=True)
        assert target.dtype == dtype and target.shape == (500,)
        assert float(torch.max(torch.abs(target - source))) == 0.0
        target.zero_()
        _add_at_indices(target, indices * 0 + 203, source, is_sorted=True)
        assert float(target[203]) == 500.0
        target[203] = 0.0
        assert float(torch.max(torch.abs(target))) == 0.0
    A = ref_vectors["add_at_indices"]
    target = torch.tensor(A["target"], device=dev)
    _add_at_indices(target, torch.tensor(A["indices"], device=dev), torch.tensor(A["source_values"], device=dev))
    assert target.dtype == torch.int64 and target.tolist() == A["expected"]


@pytest.mark.parametrize("n,m,order", [(1, 1, "random"), (63, 5, "random"), (65, 3, "sorted"),
                                       (100_003, 50, "random"), (100_003, 50, "sorted"),
                                       (1_000_000, 117_649, "random"), (200_000, 1, "random")])
def test_add_at_indices_vs_oracle(n, m, order, dev):
    from torchquad_amd.integration.utils import _add_at_indices

    rng = np.random.default_rng(n + m)
    idx = rng.integers(0, m, n)
    if order == "sorted":
        idx = np.sort(idx)
    # integers: bit-exact in every dtype (the float sums stay below 2^24)
    src = rng.integers(-8, 9, n)
    for dtype, npd in ((torch.int64, np.int64), (torch.int32, np.int32), (torch.float64, np.float64),
                       (torch.float32, np.float32)):
        t0 = rng.integers(-3, 4, m).astype(npd)
        ref = ov.add_at_indices(t0.copy(), idx, src.astype(npd))
        t = torch.tensor(t0, device=dev)
        _add_at_indices(t, torch.tensor(idx, device=dev), torch.tensor(src.astype(npd), device=dev),
                        is_sorted=order == "sorted")
        assert np.array_equal(t.cpu().numpy(), ref), dtype
    # real values: FP64 to 1e-12 of the entry's absolute sum
    f = rng.standard_normal(n)
    ref = ov.add_at_indices(np.zeros(m), idx, f)
    scale = ov.add_at_indices(np.zeros(m), idx, np.abs(f)) + 1e-300
    t = torch.zeros(m, dtype=torch.float64, device=dev)
    _add_at_indices(t, torch.tensor(idx, device=dev), torch.tensor(f, device=dev))
    assert np.max(np.abs(t.cpu().numpy() - ref) / scale) <= 1e-12


def test_add_at_indices_errors(dev):
    from torchquad_amd.integration.utils import _add_at_indices

    t = torch.zeros(4, device=dev)
    with pytest.raises(ValueError):
        _add_at_indices(t, torch.tensor([0, 1], device=dev), torch.ones(2, dtype=torch.float64, device=dev))
    with pytest.raises(ValueError):
        _add_at_indices(t, torch.tensor([0, 1], device=dev), torch.ones(3, device=dev))
    # out-of-range entries are skipped, the rest still land, and the call raises
    with pytest.raises(IndexError):
        _add_at_indices(t, torch.tensor([0, 4, -1, 3], device=dev), torch.ones(4, device=dev))
    assert t.tolist() == [1.0, 0.0, 0.0, 1.0]
    _add_at_indices(t, torch.tensor([], dtype=torch.int64, device=dev), torch.ones(0, device=dev))
    assert t.tolist() == [1.0, 0.0, 0.0, 1.0]


@pytest.mark.parametrize("N,dim", [(1500, 2), (3000, 1), (4100, 3), (64, 4), (65, 2)])
def test_map_update_many_bins_matches_oracle(N, dim, dev):
    """update_map with more bins than one block's threads (chunked block scans and run sums),
    zero-count stretches longer than the ten fill passes close, against the oracle."""
    rng = np.random.default_rng(N + dim)
    m = tq.VEGASMap(N, dim, "torch", torch.float64)
    om = ov.Map(N, dim)
    w = rng.uniform(0.0, 3.0, (dim, N)) ** 3
    c = rng.integers(0, 4, (dim, N))
    c[:, N // 3:N // 3 + min(30, N // 4)] = 0  # a run of empty bins
    w[c == 0] = 0.0
    m._dev.weights.copy_(torch.tensor(w))
    m._dev.counts.copy_(torch.tensor(c, dtype=torch.float64))
    om.weights, om.counts = w.copy(), c.copy()
    sw_dev = tq.VEGASMap._smooth_map(torch.tensor(w, device=dev), torch.tensor(c, device=dev), 0.5)
    sw_ref = ov.Map.smooth(w, c, 0.5)
    np.testing.assert_allclose(sw_dev.cpu().numpy(), sw_ref, rtol=1e-13, atol=1e-15)
    m.update_map()
    om.update_map()
    x_dev = m.x_edges.cpu().numpy()
    np.testing.assert_allclose(x_dev, om.x_edges, rtol=0, atol=1e-12)
    assert np.all(np.diff(x_dev, axis=1) > 0) and np.all(x_dev[:, 0] == 0.0) and np.all(x_dev[:, -1] == 1.0)


def test_default_bins_vegas_runs_fast(dev):
    """The reference's default map size (N_increment // 10 bins: 4000 here) on a 1e6-eval run."""
    import time

    f = tq.DoubleGaussian(6, width=0.1)
    dom = torch.tensor([[0.0, 1.0]] * 6, dtype=torch.float32, device=dev)
    v = tq.VEGAS()
    v.integrate(f, 6, N=10**6, integration_domain=dom, seed=1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    val = float(v.integrate(f, 6, N=10**6, integration_domain=dom, seed=2))
    torch.cuda.synchronize()
    assert v.map.N_intervals == 4000
    assert abs(val - 1.0) < 0.02
    assert time.perf_counter() - t0 < 1.0


def test_c3_shape_iteration_and_full_run(configs, dev, parity):
    """The VEGAS workload at its real shape: d = 6, 50 bins per axis, FP32
    samples, N_increment = 1e6 (N = 2.5e7), so N_strat = 7 and 117,649 cubes
    of 8 samples (941,192 evaluations per fresh iteration).

    One fused iteration against the oracle on the same Philox stream:
      * nh: bit-exact (8 per cube);
      * map bin counts from the fused kernel's bin indices: bit-exact
        (941,192 x 6 indices);
      * tq_vegas_map_eval indices of the same 941,192 FP32 samples: bit-exact;
      * per-cube JF / JF2 and the iteration's result: the kernel evaluates
        exp(-|x - c|^2 / w^2) in FP32 with 1/w^2 = 100, so a relative error
        of about 1e-7 in |x - c|^2 becomes up to ~1e-5 in f; the bar is
        2e-5 relative for every cube that carries >= 1e-6 of the largest
        JF, 1e-5 for the iteration result;
      * get_NH of the dh that the iteration leaves: bit-exact given that dh.
    The full run: within 3 reported standard errors of the analytic value.
    """
    c = configs["C3"]
    N, seed, d = 25_000_000, 3, 6
    f = tq.DoubleGaussian(d, width=c["width"])
    dom = torch.tensor(c["domain"], dtype=torch.float32, device=dev)
    v = tq.VEGAS()
    v._setup(f, d, N, dom, seed, None, True, 0, 0, 20, None, c["N_intervals"])
    dev_st = v._dev
    assert v.strat.N_strat == 7 and dev_st.n_cubes == 117_649 and v._starting_N == 1_000_000
    steps = v._steps(use_warmup=False)
    next(steps)  # iteration 1 (fused: samples, map, integrand, histograms, per-cube sums, update)
    torch.cuda.synchronize()
    slot = v._slots[0][0]
    nh = dev_st.nh.cpu().numpy()

    ost = ov.Strat(1_000_000, d, np.float32)
    om = ov.Map(c["N_intervals"], d, np.float32)
    onh = ost.get_NH(1_000_000)
    assert np.array_equal(nh, onh) and int(onh.sum()) == 941_192
    total = int(onh.sum())
    u = native.uniform(seed, 0, total, d, np.float32)
    y = ost.get_Y(onh, u)
    # the integrand on the inputs the kernel holds: centres rounded to FP32
    cen = (float(np.float32(1.0 / 3.0)), float(np.float32(2.0 / 3.0)))
    x = om.get_X(y).astype(np.float32).astype(np.float64)
    fv = fam.double_gaussian(x, c["width"], cen).astype(np.float32).astype(np.float64)
    jf = fv * om.get_Jac(y)
    om.accumulate_weight(y, jf**2)
    # bin counts of the fused kernel (left in the map state by the iteration's reduce)
    assert np.array_equal(dev_st.counts.cpu().numpy(), om.counts.astype(np.float64))
    np.testing.assert_allclose(dev_st.weights.cpu().numpy(), om.weights, rtol=2e-5)
    # public map evaluation on the same samples
    ids = v.map._get_interval_ID(torch.tensor(y, device=dev)).cpu().numpy()
    assert np.array_equal(ids, om.ids(y))
    JF, JF2 = ost.accumulate_weight(onh, jf)
    got_jf, got_jf2 = dev_st.jf_cube.cpu().numpy(), dev_st.jf2_cube.cpu().numpy()
    sig = JF >= 1e-6 * JF.max()
    rel = np.abs(got_jf - JF)[sig] / JF[sig]
    parity.record(float(rel.max()), 2e-5)
    assert rel.max() <= 2e-5
    assert np.all(np.abs(got_jf2 - JF2)[sig] <= 4e-5 * JF2[sig])
    assert np.all(np.abs(got_jf - JF)[~sig] <= 1e-11 * JF.max())
    ih = JF * (ost.V_cubes / onh)
    res_dev = float(dev_st.iter_stats[slot, 0])
    assert abs(res_dev - ih.sum()) <= 1e-5 * ih.sum()
    # dh of the iteration, normalised as the next get_NH does; then nh from it
    from torchquad_amd._lib import call

    call("tq_vegas_normalize_dh", dev_st.struct, slot, dev_st.stream())
    dh = dev_st.dh.cpu().numpy()
    assert abs(dh.sum() - 1.0) < 1e-12
    ost.JF, ost.JF2 = got_jf.copy(), got_jf2.copy()
    ost.update_DH()
    np.testing.assert_allclose(dh, ost.dh, rtol=1e-12, atol=1e-18)
    ost.dh = dh
    for n_exp in (1_000_000, 2_000_000, 3_141_593):
        nh_dev = v.strat.get_NH(n_exp).cpu().numpy()
        assert np.array_equal(nh_dev, ost.get_NH(n_exp)), n_exp
    assert len(np.unique(nh_dev)) > 10  # the counts are no longer uniform

    # the full run
    v = tq.VEGAS()
    val = float(v.integrate(f, d, N=N, integration_domain=dom, seed=seed, N_intervals=c["N_intervals"]))
    err = float(v._get_error())
    assert v.strat.N_strat == 7 and v.map.N_intervals == 50
    assert 0 < err < 1e-3
    assert abs(val - c["analytic"]) <= 3 * err, (val, c["analytic"], err)
