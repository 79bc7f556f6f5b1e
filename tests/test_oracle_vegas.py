"""Pins the VEGAS oracle (CPU) against the reference's vectors, invariants and
an independent formulation of the edge update."""
import math

import numpy as np
import pytest

from oracle import vegas as ov


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_interval_golden_vectors(ref_vectors, dtype):
    g = ref_vectors["vegas_map_interval"]
    m = ov.Map(g["N_intervals"], 3, dtype)
    y = np.array(g["y"], dtype=dtype)
    assert np.array_equal(m.ids(y), np.array(g["ID"]))
    assert np.max(np.abs(m.offsets(y) - np.array(g["offset"]))) <= g["offset_tol"]
    assert np.max(np.abs(m.get_X(y) - y)) <= 3e-7


def test_smooth_golden_vector(ref_vectors):
    g = ref_vectors["vegas_smooth_map"]
    out = ov.Map.smooth(g["weights"], np.array(g["counts"]), g["alpha"])
    assert np.max(np.abs(out - np.array(g["expected"]))) <= g["tol"]


def test_counts_and_update_invariants(ref_vectors):
    g = ref_vectors["vegas_map_counts"]
    dim, N = g["dim"], g["N_intervals"]
    m = ov.Map(N, dim)
    y1 = np.linspace(0.0, g["y_max"], g["N_per_dim"])
    y = np.stack([a.ravel() for a in np.meshgrid(*([y1] * dim), indexing="ij")], axis=1)
    f = np.prod(np.exp(y), axis=1)
    jac = m.get_Jac(y)
    assert np.max(np.abs(jac - 1.0)) < 1e-14
    m.accumulate_weight(y, (f * jac) ** 2)
    assert np.min(m.weights[:, 1:] - m.weights[:, :-1]) > 0.0
    assert np.all(m.counts == g["count"])
    m.update_map()
    assert np.all(m.x_edges[:, 0] == 0.0) and np.all(m.x_edges[:, -1] == 1.0)
    assert np.max(np.abs(np.sum(m.dx_edges, axis=1) - 1.0)) < 3e-7
    assert np.min(m.dx_edges) > 0.0
    assert np.max(m.dx_edges[:, 1:] - m.dx_edges[:, :-1]) < 0.0
    assert np.max(np.abs(m.get_X(y)[0])) == 0.0


def test_update_map_matches_inverse_cdf():
    rng = np.random.default_rng(0)
    for N in (2, 5, 50, 333):
        m = ov.Map(N, 2)
        # a non-uniform starting map
        m.x_edges[:, 1:-1] = np.sort(rng.uniform(0, 1, (2, N - 1)), axis=1)
        m.dx_edges = np.diff(m.x_edges, axis=1)
        m.weights = rng.uniform(0.1, 5.0, (2, N))
        m.counts = rng.integers(1, 20, (2, N))
        sw = ov.Map.smooth(m.weights, m.counts, m.alpha)
        old_x, old_dx = m.x_edges.copy(), m.dx_edges.copy()
        m.update_map()
        for i in range(2):
            ref = ov.edges_by_inverse_cdf(sw[i], old_x[i], old_dx[i])
            np.testing.assert_allclose(m.x_edges[i], ref, rtol=0, atol=1e-12)


def test_stratification_invariants(ref_vectors):
    g = ref_vectors["vegas_strat_fresh"]
    st = ov.Strat(g["N_increment"], g["dim"])
    assert st.N_strat == g["N_strat"]
    nh = st.get_NH(g["nevals_exp"])
    assert np.all(nh == g["nh"])
    from oracle import native

    u = native.uniform(0, 0, int(nh.sum()), 3, np.float64)
    y = st.get_Y(nh, u)
    assert y.shape == (nh.sum(), 3) and np.all(y >= 0) and np.all(y <= 1)
    # samples of cube h lie in its sub-cube
    dig = st.digits(np.repeat(np.arange(st.N_cubes), nh))
    assert np.all(np.floor(y * st.N_strat) == dig)
    f = np.prod(np.exp(y), axis=1)
    jf, jf2 = st.accumulate_weight(nh, f)
    assert np.min(jf2) >= 0 and np.min(jf**2 - jf2) >= 0
    st.update_DH()
    assert abs(st.dh.sum() - 1.0) < 1e-12 and st.dh[-1] > st.dh[0]
    nh2 = st.get_NH(g["nevals_exp"])
    assert nh2[-1] > nh2[0]


def test_driver_special_cases(ref_vectors):
    g = ref_vectors["vegas_special"]
    r = ov.run(lambda x: x[:, 0] * 0.0, 2, 10000, g["constant_domain"], seed=0)
    assert r["result"] == 0.0
    r = ov.run(lambda x: x[:, 0] * 0.0 + g["constant_value"], 2, 10000, g["constant_domain"], seed=0)
    assert abs(r["result"] - g["constant_integral"]) < g["constant_tol"]


def test_driver_peak_accuracy():
    from oracle import families as fam

    r = ov.run(lambda x: fam.double_gaussian(x, 0.1), 3, 30000, [[0.0, 1.0]] * 3, seed=1)
    exact = fam.double_gaussian_integral(3, [0.0] * 3, [1.0] * 3)
    assert abs(r["result"] - exact) < 0.03 * exact


def test_add_at_indices_known_answers(ref_vectors):
    # utils_integration_test.py:92-115
    A = ref_vectors["add_at_indices"]
    t = np.array(A["target"], dtype=np.int64)
    ov.add_at_indices(t, np.array(A["indices"]), np.array(A["source_values"], dtype=np.int64))
    assert t.dtype == np.int64 and t.tolist() == A["expected"]
    t = np.zeros(500)
    ov.add_at_indices(t, np.arange(500), np.ones(500))
    assert np.max(np.abs(t - 1.0)) == 0.0
    t = np.zeros(500)
    ov.add_at_indices(t, np.full(500, 203), np.ones(500))
    assert t[203] == 500.0 and np.count_nonzero(t) == 1


@pytest.mark.parametrize("n,m", [(1, 1), (1000, 7), (5000, 4096), (3000, 50)])
def test_add_at_indices_vs_numpy_add_at(n, m):
    rng = np.random.default_rng(n + m)
    idx = rng.integers(0, m, n)
    src = rng.integers(-1000, 1000, n).astype(np.int64)
    t = rng.integers(-5, 5, m).astype(np.int64)
    ref = t.copy()
    np.add.at(ref, idx, src)
    assert np.array_equal(ov.add_at_indices(t, idx, src), ref)
    # a dropped run (first or last) or a wrong run boundary fails this too
    f = rng.standard_normal(n)
    tf = np.zeros(m)
    reff = np.zeros(m)
    np.add.at(reff, idx, f)
    np.testing.assert_allclose(ov.add_at_indices(tf, idx, f), reff, rtol=1e-12, atol=1e-12)
