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


# ---- zero-count fill (vegas_map.py:118-142), hand-traced expectations ------
# Each round of the reference loop first lets every empty bin take its right
# neighbour's value (an empty bin whose right neighbour is filled becomes
# filled), then every still-empty bin its left neighbour's value.  The
# expected rows below were traced by hand through those two steps.
FILL_CASES = [
    # one empty bin between two filled ones: filled from the right
    ([[2.0, 0.0, 9.0]], [[1, 0, 3]], [[2.0, 3.0, 3.0]]),
    # two empty bins: the right one from the right, the left one from the left
    ([[4.0, 0.0, 0.0, 10.0]], [[2, 0, 0, 5]], [[2.0, 2.0, 2.0, 2.0]]),
    ([[4.0, 0.0, 0.0, 30.0]], [[1, 0, 0, 5]], [[4.0, 4.0, 6.0, 6.0]]),
    # three empty bins: the middle one is reached from the right in round two
    ([[1.0, 0.0, 0.0, 0.0, 8.0]], [[1, 0, 0, 0, 1]], [[1.0, 1.0, 8.0, 8.0, 8.0]]),
    # empty bins at the right edge: only the left fill reaches them
    ([[3.0, 5.0, 0.0, 0.0]], [[1, 1, 0, 0]], [[3.0, 5.0, 5.0, 5.0]]),
    # empty bins at the left edge: only the right fill reaches them
    ([[0.0, 0.0, 7.0, 1.0]], [[0, 0, 1, 1]], [[7.0, 7.0, 7.0, 1.0]]),
    # two axes are filled independently
    ([[0.0, 6.0, 0.0], [5.0, 0.0, 0.0]], [[0, 2, 0], [1, 0, 0]], [[3.0, 3.0, 3.0], [5.0, 5.0, 5.0]]),
]


@pytest.mark.parametrize("case", range(len(FILL_CASES)))
def test_zero_count_fill_hand_traced(case):
    w, c, expected = FILL_CASES[case]
    out = ov.Map.fill_zero_counts(np.array(w), np.array(c))
    assert out.tolist() == expected


def test_zero_count_fill_long_gap():
    # a gap of 22 empty bins closes by one bin per side per round; after the
    # ten rounds bins 11 and 12 are still empty and keep weight zero
    n = 25
    w = np.zeros((1, n))
    c = np.zeros((1, n), dtype=np.int64)
    w[0, 0], c[0, 0] = 2.0, 1
    w[0, 23], c[0, 23] = 9.0, 3
    w[0, 24], c[0, 24] = 5.0, 1
    out = ov.Map.fill_zero_counts(w, c)[0]
    assert out[:11].tolist() == [2.0] * 11
    assert out[11:13].tolist() == [0.0, 0.0]
    assert out[13:24].tolist() == [3.0] * 11
    assert out[24] == 5.0
    # the smoothed map then has non-zero entries next to the remaining gap
    # ([1,6,1] of (2, 0, 0) and (0, 0, 3)) and none inside a wider one
    d = ov.Map.smooth(w, c, 0.5)[0]
    assert d[11] > 0 and d[12] > 0
    w2 = np.zeros((1, 30))
    c2 = np.zeros((1, 30), dtype=np.int64)
    w2[0, 0], c2[0, 0] = 1.0, 1
    w2[0, 29], c2[0, 29] = 1.0, 1
    filled = ov.Map.fill_zero_counts(w2, c2)[0]
    assert np.count_nonzero(filled == 0.0) == 8
    d2 = ov.Map.smooth(w2, c2, 0.5)[0]
    # bins 12..17 are empty with empty neighbours: smoothed weight stays 0
    assert d2[12:18].tolist() == [0.0] * 6 and np.all(d2[:11] > 0) and np.all(d2[19:] > 0)


def _compressed(raw):
    out = np.zeros_like(raw)
    nz = raw != 0
    out[nz] = (raw[nz] - 1.0) / np.log(raw[nz])
    return out


def test_smooth_closed_form():
    # uniform averages: every smoothed entry is 1/n before compression, and
    # ((1/n - 1) / ln(1/n))^alpha after it
    for n, alpha in ((4, 0.5), (50, 0.5), (7, 1.0)):
        d = ov.Map.smooth(np.full((2, n), 3.0), np.full((2, n), 3), alpha)
        expect = ((1.0 / n - 1.0) / math.log(1.0 / n)) ** alpha
        np.testing.assert_allclose(d, expect, rtol=1e-15, atol=0)
    # a single non-zero bin: [7,1] at the left edge, [1,6,1] inside, [1,7] at
    # the right edge, all over 8 * sum
    w = np.array([[0.0, 0.0, 8.0, 0.0, 0.0]])
    c = np.ones((1, 5), dtype=np.int64)
    d = ov.Map.smooth(w, c, 1.0)[0]
    raw = np.array([0.0, 8.0, 48.0, 8.0, 0.0]) / 64.0
    expect = _compressed(raw)
    np.testing.assert_allclose(d, expect, rtol=1e-15, atol=0)
    w = np.array([[8.0, 0.0, 0.0, 0.0, 16.0]])
    d = ov.Map.smooth(w, c, 1.0)[0]
    raw = np.array([56.0, 8.0, 0.0, 16.0, 112.0]) / 192.0
    expect = _compressed(raw)
    np.testing.assert_allclose(d, expect, rtol=1e-15, atol=0)
    # a row that sums to zero cannot be smoothed
    assert ov.Map.smooth(np.zeros((1, 4)), np.zeros((1, 4), dtype=np.int64), 0.5) is None


# ---- driver estimators and abort schedule (vegas.py:161-209, 318-362) ------
def test_driver_estimators_closed_form():
    # weights 1, 1/4, 1/16: result (1 + 2/4 + 4/16) / (21/16) = 4/3,
    # error 1/sqrt(21/16), chi^2 = (1/3)^2 + (2/3)^2/4 + (8/3)^2/16 = 2/3
    r, s = [1.0, 2.0, 4.0], [1.0, 4.0, 16.0]
    assert ov._result(r, s) == pytest.approx(4.0 / 3.0, rel=1e-15)
    assert ov._error(s) == pytest.approx(4.0 / math.sqrt(21.0), rel=1e-15)
    assert ov._chisq(r, s) == pytest.approx(2.0 / 3.0, rel=1e-14)
    # a zero variance: plain mean, zero skipped in the error, inf in chi^2
    r, s = [1.0, 2.0, 6.0], [0.5, 0.0, 2.0]
    assert ov._result(r, s) == 3.0
    assert ov._error(s) == pytest.approx(1.0 / math.sqrt(2.5), rel=1e-15)
    assert ov._chisq(r, s) == math.inf
    # every variance zero: the error is sigma2[0]
    assert ov._error([0.0, 0.0]) == 0.0
    # entries equal to the result are left out of chi^2 (even with zero variance)
    assert ov._chisq([2.0, 2.0], [0.0, 0.0]) == 0.0


def _abort(it, r, s, eps_rel=0.0, eps_abs=0.0, sN=1000, inc=1000, fevals=0, N=10**9, max_it=20):
    return ov.abort_check(it, r, s, eps_rel, eps_abs, sN, inc, fevals, N, max_it)


def test_abort_check_branches():
    ok = [1.0] * 5
    tiny = [0.01] * 5  # error 1/sqrt(500) = 0.0447, chi^2 = 0
    # accuracy reached
    assert _abort(5, ok, tiny, eps_rel=0.1) == (True, 1000)
    assert _abort(5, ok, tiny, eps_abs=0.05) == (True, 1000)
    # accuracy not reached, stable: N grows by the increment (the sqrt bound
    # is far above it with eps_rel = 0)
    assert _abort(5, ok, tiny) == (False, 2000)
    # stable with eps_rel: N * sqrt(acc / (eps_rel + 1e-8)), acc = 0.0121,
    # eps_rel = 0.01: 10000 * 1.09999945 -> 10999 (11000 without the 1e-8)
    sig = [5 * 0.0121**2] * 5
    assert _abort(5, ok, sig, eps_rel=0.01, sN=10000, inc=10000) == (False, 10999)
    # a zero result adds the increment
    assert _abort(5, [0.0] * 5, tiny) == (False, 2000)
    # unstable: results 3,-1,3,-1,1 with variance 2 give I = 1 and
    # chi^2 = (4 + 4 + 4 + 4 + 0) / 2 = 8 > 5: increment
    assert _abort(5, [3.0, -1.0, 3.0, -1.0, 1.0], [2.0] * 5) == (False, 2000)
    # chi^2 / 5 exactly 1 (results 3,-1,2,0,1, variance 2: (4+4+1+1)/2 = 5):
    # neither branch, N unchanged
    assert _abort(5, [3.0, -1.0, 2.0, 0.0, 1.0], [2.0] * 5) == (False, 1000)
    # a zero variance that differs from the mean: chi^2 = inf, increment
    assert _abort(5, [1.0, 2.0, 1.0, 1.0, 1.0], [1.0, 0.0, 1.0, 1.0, 1.0]) == (False, 2000)
    # budget: fevals + 5 * new N > N stops, == N goes on
    assert _abort(5, ok, tiny, fevals=5000, N=15000) == (False, 2000)
    assert _abort(5, ok, tiny, fevals=5001, N=15000) == (True, 2000)
    # iteration cap: it + 5 > max_iterations stops
    assert _abort(15, ok, tiny, max_it=20) == (False, 2000)
    assert _abort(20, ok, tiny, max_it=20) == (True, 2000)
    assert _abort(15, ok, tiny, max_it=19) == (True, 2000)


def test_abort_schedule_scripted_trace():
    # A scripted run of the driver loop (vegas.py:130-146) with mocked
    # iteration results: each iteration uses starting_N evaluations, the
    # results of a block of five are noisy (chi^2 / 5 = 8 / 5 > 1) for the
    # first block and then stable.  N = 230000, max_iterations = 20 give
    # starting_N = N_increment = 230000 // 25 = 9200 (vegas.py:114-115).
    N, max_it = 230000, 20
    sN = inc = N // (max_it + 5)
    assert sN == 9200
    blocks = [([3.0, -1.0, 3.0, -1.0, 1.0], [2.0] * 5), ([1.0] * 5, [0.01] * 5), ([1.0] * 5, [0.01] * 5)]
    fevals, it, trace = 0, 0, []
    while True:
        it += 1
        fevals += sN
        if it % 5 > 0:
            continue
        r, s = blocks[it // 5 - 1]
        stop, sN = ov.abort_check(it, r, s, 0.0, 0.0, sN, inc, fevals, N, max_it)
        trace.append((it, fevals, sN, stop))
        if stop:
            break
    # block 1: unstable, N 9200 -> 18400; 46000 + 92000 <= 230000 goes on.
    # block 2: stable, N 18400 -> 27600; 138000 + 138000 > 230000 stops.
    assert trace == [(5, 46000, 18400, False), (10, 138000, 27600, True)]
