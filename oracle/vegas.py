"""VEGAS restatement in numpy (oracle).

Follows torchquad/integration:
  vegas_map.py:15-42     uniform edges, dx = 1/N, zero weights/counts
  vegas_map.py:44-111    get_X, get_Jac, interval id / offset, accumulate_weight
  vegas_map.py:113-172   _smooth_map (zero-count fill, [1,6,1]/8, range compression)
  vegas_map.py:185-261   update_map (fp64 cumsum, add-at, edge interpolation, repair)
  vegas_stratification.py:13-165   N_strat, get_NH, get_Y, accumulate_weight, update_DH
  vegas.py:30-362        driver: budget split, warmup, iterations, abort rules, EQ 30-32
  utils.py:61-107        _add_at_indices (add_at_indices below)

State is FP64 and samples use `dtype`, the precisions of the device
implementation.  Random numbers come from the library's documented Philox
stream (oracle.native.uniform): warmup iteration w uses N_samples counters,
main iteration i reserves n_exp + 3*N_cubes + 1024 counters and sample s of
the iteration uses counter base + s.

Pinned by: the reference's golden vectors (vegas_map_test.py:28-48, 84-100),
its invariants (vegas_map_test.py:55-81, 104-128; vegas_stratification_test.py
:24-68) and special cases (vegas_test.py:164-184), and by an independent
inverse-CDF formulation of the edge update (tests/test_oracle_vegas.py).
"""
import math

import numpy as np

from . import native

SB = 1024


def add_at_indices(target, indices, source):
    """target[indices[i]] += source[i] in place (torchquad/integration/utils.py:61-107).

    Follows the reference's numpy branch (:78-107): order the pairs by index,
    sum every run of equal indices, add each run's sum to its entry.  Pinned by
    the reference's known answer (tests/utils_integration_test.py:109-115, in
    tests/golden/reference_vectors.json), its identity and single-entry cases
    (:92-104) and numpy's np.add.at on random inputs (tests/test_oracle.py).
    """
    idx = np.asarray(indices)
    if idx.size == 0:
        return target
    order = np.argsort(idx, kind="stable")
    idx = idx[order]
    src = np.asarray(source)[order]
    starts = np.flatnonzero(np.r_[True, idx[1:] != idx[:-1]])
    target[idx[starts]] += np.add.reduceat(src, starts).astype(target.dtype)
    return target


def n_strat_for(N_increment, dim):
    n = int((N_increment / 4.0) ** (1.0 / dim))
    return 1000 if n > 1000 else n


class Map:
    def __init__(self, N_intervals, dim, dtype=np.float64, alpha=0.5):
        self.N = N_intervals
        self.dim = dim
        self.dtype = dtype
        self.alpha = alpha
        step = 1.0 / N_intervals
        row = np.array([j * step for j in range(N_intervals)] + [1.0])
        self.x_edges = np.tile(row, (dim, 1))
        self.dx_edges = np.full((dim, N_intervals), 1.0 / N_intervals)
        self.reset()

    def reset(self):
        self.weights = np.zeros((self.dim, self.N))
        self.counts = np.zeros((self.dim, self.N), dtype=np.int64)

    def ids(self, y):
        return np.floor(y * self.dtype(self.N)).astype(np.int64)

    def offsets(self, y):
        yn = y * self.dtype(self.N)
        return yn - np.floor(yn)

    def get_X(self, y):
        ID, off = self.ids(y), self.offsets(y)
        res = np.zeros(y.shape)
        for i in range(self.dim):
            res[:, i] = self.x_edges[i, ID[:, i]] + self.dx_edges[i, ID[:, i]] * off[:, i].astype(np.float64)
        return res

    def get_Jac(self, y):
        ID = self.ids(y)
        jac = np.ones(y.shape[0])
        for i in range(self.dim):
            jac *= self.N * self.dx_edges[i][ID[:, i]]
        return jac

    def accumulate_weight(self, y, jf2):
        ID = self.ids(y)
        for i in range(self.dim):
            np.add.at(self.weights[i], ID[:, i], jf2)
            np.add.at(self.counts[i], ID[:, i], 1)

    @staticmethod
    def fill_zero_counts(weights, counts):
        """Average weight per bin, bins without samples filled from neighbours
        (vegas_map.py:118-142): divide by the count where it is non-zero, then
        up to ten rounds of: every empty bin takes its right neighbour's value
        and stops being empty if that neighbour was not; then every empty bin
        takes its left neighbour's value likewise.  Each round closes a gap by
        one bin from each side; bins more than ten from either side of a gap
        keep weight zero."""
        weights = np.array(weights, dtype=np.float64)
        counts = np.asarray(counts)
        z_idx = counts == 0
        if np.any(z_idx):
            nnz = np.logical_not(z_idx)
            weights[nnz] /= counts[nnz]
            for _ in range(10):
                weights[:, :-1] = np.where(z_idx[:, :-1], weights[:, 1:], weights[:, :-1])
                z_idx[:, :-1] = z_idx[:, :-1] * z_idx[:, 1:]
                weights[:, 1:] = np.where(z_idx[:, 1:], weights[:, :-1], weights[:, 1:])
                z_idx[:, 1:] = z_idx[:, 1:] * z_idx[:, :-1]
                if not np.any(z_idx):
                    break
        else:
            weights /= counts
        return weights

    @staticmethod
    def smooth(weights, counts, alpha):
        """_smooth_map (vegas_map.py:113-172): zero-count fill, [7,1] / [1,6,1]
        / [1,7] smoothing over 8 times the row sum, then range compression
        ((d - 1) / ln d)^alpha of the non-zero entries.  None when a row sums
        to zero."""
        weights = Map.fill_zero_counts(weights, counts)
        dim, n = weights.shape
        sums = np.reshape(np.sum(weights, axis=1), [dim, 1])
        if np.any(sums == 0.0):
            return None
        i = n - 2
        d = np.concatenate(
            [7.0 * weights[:, 0:1] + weights[:, 1:2], weights[:, :-2] + 6.0 * weights[:, 1:-1] + weights[:, 2:],
             weights[:, i:i + 1] + 7.0 * weights[:, i + 1:i + 2]], axis=1)
        d = d / (8.0 * sums)
        nz = d != 0
        with np.errstate(divide="ignore", invalid="ignore"):
            d[nz] = ((d[nz] - 1.0) / np.log(d[nz])) ** alpha
        return d

    def update_map(self):
        """Returns (updated, n_nonfinite_axes)."""
        sw = self.smooth(self.weights, self.counts, self.alpha)
        if sw is None:
            self.reset()
            return False, 0
        delta = np.sum(sw, axis=1) / self.N
        bad = 0
        for i in range(self.dim):
            mult = (np.cumsum(sw[i, :-1].astype(np.float64)) / delta[i]).astype(np.int64)
            num = np.zeros(self.N + 1, dtype=np.int64)
            np.add.at(num, mult, 1)
            val = np.zeros(self.N + 1)
            np.add.at(val, mult, sw[i, : self.N - 1])
            idx = np.cumsum(num[:-2])
            accu = np.cumsum(delta[i] - val[:-2])
            self.x_edges[i][1:-1] = self.x_edges[i][idx] + accu / sw[i][idx] * self.dx_edges[i][idx]
            fin = np.isfinite(self.x_edges[i])
            if not np.all(fin):
                bad += 1
                mid = 0.5 * (self.x_edges[i][:-2] + self.x_edges[i][2:])
                self.x_edges[i][1:-1] = np.where(fin[1:-1], self.x_edges[i][1:-1], mid)
                if not np.all(np.isfinite(self.x_edges[i])):
                    raise RuntimeError("Could not replace all infinite edges")
            self.dx_edges[i] = self.x_edges[i][1:] - self.x_edges[i][:-1]
        self.reset()
        return True, bad


class Strat:
    def __init__(self, N_increment, dim, dtype=np.float64, beta=0.75):
        self.dim = dim
        self.N_strat = n_strat_for(N_increment, dim)
        self.beta = beta
        self.N_cubes = self.N_strat**dim
        self.V_cubes = (1.0 / self.N_strat) ** dim
        self.dtype = dtype
        self.dh = np.ones(self.N_cubes) * 1.0 / self.N_cubes
        self.JF = np.zeros(self.N_cubes)
        self.JF2 = np.zeros(self.N_cubes)
        self.strat_counts = np.zeros(self.N_cubes)

    def get_NH(self, nevals_exp):
        nh = np.floor(self.dh * nevals_exp)
        return np.clip(nh, 2, None).astype(np.int64)

    def digits(self, idx):
        pts = idx.reshape(-1, 1) // (self.N_strat ** np.arange(self.dim))
        pts[:, :-1] = pts[:, :-1] - self.N_strat * pts[:, 1:]
        return pts

    def get_Y(self, nevals, u):
        pos = self.digits(np.arange(len(nevals)))
        pos = np.repeat(pos, nevals, axis=0).astype(self.dtype)
        y = (pos + u.astype(self.dtype)) / self.dtype(self.N_strat)
        y[y >= 1.0] = self.dtype(0.999999)
        return y

    def accumulate_weight(self, nevals, w):
        idx = np.repeat(np.arange(self.N_cubes), nevals)
        self.JF = np.zeros(self.N_cubes)
        self.JF2 = np.zeros(self.N_cubes)
        np.add.at(self.JF, idx, w)
        np.add.at(self.JF2, idx, w**2.0)
        self.strat_counts = nevals.astype(np.float64)
        return self.JF, self.JF2

    def update_DH(self):
        V2 = self.V_cubes * self.V_cubes
        d = V2 * self.JF2 / self.strat_counts - (self.V_cubes * self.JF / self.strat_counts) ** 2
        d[d < 0.0] = 0.0
        self.dh = d**self.beta
        s = np.sum(self.dh)
        if s != 0:
            self.dh = self.dh / s


def capacity_for(n_exp, n_cubes):
    return int(n_exp) + 3 * n_cubes + SB


def run(f, dim, N, domain, seed, dtype=np.float64, N_intervals=None, max_iterations=20, use_warmup=True,
        use_grid_improve=True, eps_rel=0.0, eps_abs=0.0):
    """The reference driver on the library stream.  f: [n, dim] float64 -> [n].

    Returns dict(result, results, sigma2, fevals, map, strat).
    """
    domain = np.asarray(domain, dtype=np.float64)
    starts = domain[:, 0].astype(dtype)
    sizes = (domain[:, 1] - domain[:, 0]).astype(dtype)
    vol = dtype(np.prod(sizes))
    starting_N = N // (max_iterations + 5)
    N_inc = starting_N
    if N_intervals is None:
        N_intervals = max(2, N_inc // 10)
    m = Map(N_intervals, dim, dtype)
    st = Strat(N_inc, dim, dtype)
    offset = 0
    fevals = 0

    def evaluate(y):
        xm = m.get_X(y)
        x = xm.astype(dtype) * sizes + starts
        fv = (np.asarray(f(x.astype(np.float64))).astype(dtype) * vol).astype(np.float64)
        return fv * m.get_Jac(y)

    if use_warmup:
        ns = starting_N // 5
        for _ in range(5):
            u = native.uniform(seed, offset, ns, dim, dtype)
            offset += ns
            y = u * dtype(0.999999)
            jf = evaluate(y) if ns else np.zeros(0)
            fevals += ns
            m.accumulate_weight(y, jf**2)
            m.update_map()
    results, sigma2 = [], []
    all_results = []
    it = 0
    while True:
        it += 1
        nh = st.get_NH(starting_N)
        total = int(nh.sum())
        u = native.uniform(seed, offset, total, dim, dtype)
        offset += capacity_for(starting_N, st.N_cubes)
        y = st.get_Y(nh, u)
        jf = evaluate(y)
        fevals += total
        if use_grid_improve:
            m.accumulate_weight(y, jf**2)
        JF, JF2 = st.accumulate_weight(nh, jf)
        ninv = 1.0 / nh.astype(np.float64)
        ih = JF * (ninv * st.V_cubes)
        sig2 = np.abs(JF2 * ninv * (st.V_cubes**2) - ih**2)
        results.append(float(ih.sum()))
        sigma2.append(float((sig2 * ninv).sum()))
        all_results.append((results[-1], sigma2[-1]))
        if use_grid_improve:
            m.update_map()
        st.update_DH()
        if it % 5 > 0:
            continue
        stop, starting_N = abort_check(it, results, sigma2, eps_rel, eps_abs, starting_N, N_inc, fevals, N,
                                       max_iterations)
        if stop:
            break
        results, sigma2 = [], []
    return dict(result=_result(results, sigma2), results=results, sigma2=sigma2, all=all_results, fevals=fevals,
                map=m, strat=st, iterations=it)


def abort_check(it, results, sigma2, eps_rel, eps_abs, starting_N, N_increment, fevals, N, max_iterations):
    """VEGAS._check_abort_conditions (vegas.py:161-209) as a pure function of
    the driver state.  Returns (stop, starting_N); the caller clears results
    and sigma2 when it goes on.  Called only when it % 5 == 0 (:171-173)."""
    res_abs = abs(_result(results, sigma2))
    err = _error(sigma2)
    chi2 = _chisq(results, sigma2)
    if (err <= eps_rel * res_abs or err <= eps_abs) and chi2 / 5.0 < 1.0:
        return True, starting_N
    if chi2 / 5.0 < 1.0:
        if res_abs == 0.0:
            starting_N += N_increment
        else:
            acc = err / res_abs
            starting_N = min(starting_N + N_increment, int(starting_N * math.sqrt(acc / (eps_rel + 1e-8))))
    elif chi2 / 5.0 > 1.0:
        starting_N += N_increment
    if fevals + starting_N * 5 > N:
        return True, starting_N
    if it + 5 > max_iterations:
        return True, starting_N
    return False, starting_N


def _result(results, sigma2):
    """EQ 30 (vegas.py:318-335): inverse-variance weighted mean, or the plain
    mean when any variance is zero."""
    if any(s == 0.0 for s in sigma2):
        return sum(results) / len(results)
    return sum(r / s for r, s in zip(results, sigma2)) / sum(1.0 / s for s in sigma2)


def _error(sigma2):
    """EQ 31 (vegas.py:337-346): zero variances are skipped; sigma2[0] when
    all are zero."""
    res = sum(1.0 / s for s in sigma2 if s != 0.0)
    return sigma2[0] if res == 0 else 1.0 / math.sqrt(res)


def _chisq(results, sigma2):
    """EQ 32 (vegas.py:348-362) over the entries that differ from the result.
    The reference divides backend tensors, so a zero variance there gives
    inf, not an exception."""
    I = _result(results, sigma2)
    total = 0.0
    for r, s in zip(results, sigma2):
        if r != I:
            total += (r - I) ** 2 / s if s != 0.0 else math.inf
    return total


def edges_by_inverse_cdf(sw, x_edges, dx_edges):
    """Independent formulation of the edge update: the new inner edge k is
    where the piecewise-linear cumulative of the smoothed weights (density
    sw_j / dx_j on bin j) reaches k * sum(sw) / N."""
    n = len(sw)
    total = float(np.sum(sw))
    cum = np.concatenate([[0.0], np.cumsum(sw)])
    out = np.array(x_edges, dtype=np.float64).copy()
    for k in range(1, n):
        target = k * total / n
        j = int(np.searchsorted(cum, target, side="right") - 1)
        j = min(max(j, 0), n - 1)
        out[k] = x_edges[j] + (target - cum[j]) / sw[j] * dx_edges[j]
    return out
