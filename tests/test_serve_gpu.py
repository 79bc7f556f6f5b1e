"""Serving loop: results alone vs in the mix, HBM-budget admission,
cancellation, threaded submission, VEGAS state across scheduling rounds."""
import threading

import numpy as np
import pytest
import torch

import torchquad_amd as tq
from torchquad_amd import _lib
from torchquad_amd.batch import integrate_batch
from torchquad_amd.serve import AdmissionError, CancelledError, IntegrationServer, latency_stats, serve_open_loop
from torchquad_amd.workloads import genz_batch

pytestmark = pytest.mark.gpu

TILE = 4096


def grid_submitter(it):
    return lambda s, arrival: s.submit_grid(it["rule"], it["fn"], it["dim"], N=it["N"], domain=it["domain"],
                                            arrival=arrival)


def vegas_submitter(seed):
    f = tq.DoubleGaussian(3, width=0.2)
    return lambda s, arrival: s.submit_vegas(f, 3, N=60_000, domain=[[0.0, 1.0]] * 3, seed=seed, N_intervals=20,
                                             arrival=arrival)


def test_results_same_alone_and_in_mix(dev):
    items = genz_batch(160, seed=11)
    reqs = [grid_submitter(it) for it in items]
    # three VEGAS requests interleaved with the grid requests
    for k, seed in enumerate((3, 4, 5)):
        reqs.insert(40 * (k + 1), vegas_submitter(seed))
    server = IntegrationServer(dev, dtype=torch.float32, tile_points=TILE)
    tickets = serve_open_loop(server, reqs, rate=20_000, seed=1)
    mixed = [t.result(timeout=60) for t in tickets]
    assert server.batches_launched > 1  # requests were batched continuously, not all at once

    # each request alone: one at a time on an idle server
    alone = []
    s1 = IntegrationServer(dev, dtype=torch.float32, tile_points=TILE)
    for r in reqs:
        t = r(s1, None)
        s1.run_until_idle(timeout=60)
        alone.append(t.result())
    assert mixed == alone  # bit-identical

    # and the grid results agree with a whole-batch run with the same tiles
    grid_vals = [v for v, t in zip(mixed, tickets) if t.kind == "grid"]
    ref = integrate_batch(items, dtype=torch.float32, device=dev, tile_points=TILE).tolist()
    assert grid_vals == ref


def test_grid_results_vs_plain_integrators(dev):
    items = genz_batch(40, seed=5)
    server = IntegrationServer(dev, dtype=torch.float64, tile_points=TILE)
    tickets = [grid_submitter(it)(server, None) for it in items]
    server.run_until_idle(timeout=60)
    integ = {"trapezoid": tq.Trapezoid(), "simpson": tq.Simpson(), "gauss_legendre": tq.GaussLegendre()}
    for it, t in zip(items, tickets):
        dom = torch.tensor(it["domain"], dtype=torch.float64, device=dev)
        ref = float(integ[it["rule"]].integrate(it["fn"], it["dim"], N=it["N"], integration_domain=dom))
        # same rule and points, different split into partial sums: FP64 bar
        assert abs(t.result() - ref) <= 1e-12 * abs(ref) + 1e-14, (it["rule"], it["dim"], t.result(), ref)


def test_vegas_request_matches_integrator(dev):
    f = tq.DoubleGaussian(3, width=0.2)
    server = IntegrationServer(dev, dtype=torch.float64, tile_points=TILE)
    t = server.submit_vegas(f, 3, N=60_000, domain=[[0.0, 1.0]] * 3, seed=7, N_intervals=20)
    server.run_until_idle(timeout=60)
    dom = torch.tensor([[0.0, 1.0]] * 3, dtype=torch.float64, device=dev)
    ref = float(tq.VEGAS().integrate(f, 3, N=60_000, integration_domain=dom, seed=7, N_intervals=20))
    assert t.result() == ref


def test_hbm_budget_admission(dev):
    items = genz_batch(64, seed=2)
    probe = IntegrationServer(dev, tile_points=TILE)
    sizes = [grid_submitter(it)(probe, None).nbytes for it in items]
    probe.run_until_idle(timeout=60)
    # every request fits alone, the whole set does not
    budget = max(max(sizes), sum(sizes) // 3)
    server = IntegrationServer(dev, tile_points=TILE, hbm_budget=budget)
    tickets = [grid_submitter(it)(server, None) for it in items]
    # a request that can never fit is rejected at once
    big = server.submit_grid("trapezoid", tq.GaussianPeak([1.0] * 6, [0.5] * 6), 6, n=200,
                             domain=[[0.0, 1.0]] * 6)
    assert big.done()
    with pytest.raises(AdmissionError):
        big.result()
    server.run_until_idle(timeout=60)
    assert server.peak_reserved <= budget
    assert server.batches_launched > 1  # the budget split the work
    assert server.reserved == 0
    ref = integrate_batch(items, dtype=torch.float32, device=dev, tile_points=TILE).tolist()
    assert [t.result() for t in tickets] == ref


def test_cancellation(dev):
    items = genz_batch(20, seed=3)
    server = IntegrationServer(dev, tile_points=TILE)
    tickets = [grid_submitter(it)(server, None) for it in items]
    vt = vegas_submitter(9)(server, None)
    assert server.cancel(tickets[3])
    assert server.cancel(tickets[7])
    server.step()  # admits and launches the rest; VEGAS starts
    server.cancel(vt)
    server.run_until_idle(timeout=60)
    for i, t in enumerate(tickets):
        if i in (3, 7):
            with pytest.raises(CancelledError):
                t.result()
        else:
            assert np.isfinite(t.result())
    with pytest.raises(CancelledError):
        vt.result()
    assert server.reserved == 0


def test_threaded_clients(dev):
    items = genz_batch(96, seed=4)
    results = [None] * len(items)
    with IntegrationServer(dev, tile_points=TILE) as server:

        def client(lo, hi):
            ts = [(i, grid_submitter(items[i])(server, None)) for i in range(lo, hi)]
            for i, t in ts:
                results[i] = t.result(timeout=60)

        th = [threading.Thread(target=client, args=(k * 24, (k + 1) * 24)) for k in range(4)]
        for t in th:
            t.start()
        for t in th:
            t.join()
    ref = integrate_batch(items, dtype=torch.float32, device=dev, tile_points=TILE).tolist()
    assert results == ref
    st = latency_stats(server.completed)
    assert st["count"] == len(items) and st["p99_ms"] >= st["p50_ms"] > 0


def test_gl_table_arena_clears(dev):
    """A table arena too small for the tables of one round (3000 entries for
    6733) is cleared and refilled between batches and still serves every
    request with the whole-batch values."""
    ns = [33, 2000, 33, 2500, 1500, 33, 2000, 700, 33]
    f = tq.GaussianPeak([3.0], [0.4])
    items = [{"rule": "gauss_legendre", "fn": f, "dim": 1, "n": n, "domain": [[0.0, 1.0]]} for n in ns]
    # FP64 nodes and weights + FP32 pairs: 24 B per entry
    server = IntegrationServer(dev, tile_points=TILE, gl_cache_limit=24 * 3000)
    ref = integrate_batch(items, dtype=torch.float32, device=dev, tile_points=TILE).tolist()
    for rounds in range(3):
        tickets = [server.submit_grid(it["rule"], it["fn"], 1, n=it["n"], domain=it["domain"]) for it in items]
        server.run_until_idle(timeout=60)
        assert [t.result() for t in tickets] == ref
        assert server._arena.used <= 3000 and server._arena.cap <= 3000
    many = server.submit_many(_lib.RULE_GAUSS_LEGENDRE, _lib.FAM_GAUSSIAN_PEAK, 1, [[0.0, 1.0]], [3.0], [0.4], n=ns)
    server.run_until_idle(timeout=60)
    assert many.result().tolist() == ref


def test_submit_many_matches_single_requests(dev):
    """Packed arrays (submit_many) give the same values as one request per
    integral and as a whole-batch run, bit for bit, also split into several
    arrays that share batches with single requests."""
    from torchquad_amd.workloads import genz_arrays

    count = 300
    items = genz_batch(count, seed=13)
    arr = genz_arrays(count, seed=13)
    ref = integrate_batch(items, dtype=torch.float32, device=dev, tile_points=TILE).tolist()
    server = IntegrationServer(dev, dtype=torch.float32, tile_points=TILE)
    t_all = server.submit_many(arr["rule"], arr["family"], arr["dim"], arr["domain"], arr["a"], arr["u"], arr["c"],
                               N=arr["N"])
    server.run_until_idle(timeout=60)
    vals = t_all.result()
    assert isinstance(vals, np.ndarray) and vals.shape == (count,)
    assert vals.tolist() == ref
    # three arrays and a few single requests in flight together
    parts = [slice(0, 100), slice(100, 180), slice(180, 300)]
    tickets = []
    for k, sl in enumerate(parts):
        tickets.append(server.submit_many(arr["rule"][sl], arr["family"][sl], arr["dim"][sl], arr["domain"][sl],
                                          arr["a"][sl], arr["u"][sl], arr["c"][sl], N=arr["N"][sl]))
        tickets.append(grid_submitter(items[7 * k])(server, None))
    server.run_until_idle(timeout=60)
    got = np.concatenate([tickets[0].result(), tickets[2].result(), tickets[4].result()])
    assert got.tolist() == ref
    assert [tickets[1].result(), tickets[3].result(), tickets[5].result()] == [ref[0], ref[7], ref[14]]
    st = latency_stats([t_all])
    assert st["count"] == count
    assert server.reserved == 0
