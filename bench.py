"""Benchmark of the integration hot path on one MI355X.

Workload (the configuration the headline metric is quoted on): plain
single-integral GPU path, Boole rule, dim=5 on [-1,1]^5, fp32, Gaussian peak
exp(-sum a_i^2 (x_i-u_i)^2) with a_i~U(1,5), u_i~U(-0.5,0.5) (numpy PCG64,
seed 0).  The metric is "integrand evals/s at 1e-6 relative error", so the
timed step is one Boole().integrate call at n=65 points per axis (65^5 =
1.16e9 evals), the smallest Boole-valid n whose result is within 1e-6 of the
analytic value for these parameters; n=17 (the literal 17^5 grid) is
reported as per-call latency p50/p99.  Reported beside it:
  * Monte Carlo dim=8, 1e8 samples, oscillatory integrand;
  * VEGAS dim=6, 50 bins/axis, 20 iterations x 1e6 evals, double Gaussian;
  * batched form: 8192 seeded Genz integrals (dim 1..6, N 1e3..1e6,
    trapezoid/simpson/gauss-legendre) in one graph-replayed persistent
    kernel, against a Python loop of plain per-integral calls;
  * differentiation: 1024 Gauss-Legendre dim=4 n=32 Gaussian-peak integrals
    with the hand-written backward, against a loop of per-vector autograd
    calls through the materialised-points path;
  * serving: Genz integral requests arriving as a Poisson stream at
    2000/s into the continuous-batching server (p50/p99 latency from
    arrival to result), and the same requests submitted at once;
  * the reference's CPU-runnable case (Trapezoid / Simpson, dim=3,
    101^3 points, fp64, exp(x1+x2+x3)) on the GPU, and the oracle's CPU
    rates for it and for bounded samples of the MC and VEGAS configs.

Prints one JSON line.  --gpus N>1 exits: this build measures one GPU.
"""
import argparse
import json
import math
import os
import platform
import statistics
import sys
import time
import warnings
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK_FP32_TFLOPS = 157.3  # MI355X FP32 vector (packed counted)
PEAK_HBM_GBS = 8000.0
CLOCK_GHZ = 2.4  # MI355X peak engine clock
# Algorithmic FLOPs per evaluation of the dim-5 Gaussian peak
# exp(-sum_i a_i^2 (x_i - u_i)^2) with its quadrature weight: 4d+3 arithmetic
# (+, -, x = 1, FMA = 2) plus 1 exp = 24.  The fused kernel executes far
# fewer per point (row and lane partial sums are shared across points), so
# this count is reported only as an "algorithmic-equivalent" rate; the
# roofline uses the executed FP32 operations from the committed counter pass.
DIM_C2A = 5
FLOPS_PER_EVAL_GRID_GAUSSIAN = 4 * DIM_C2A + 3 + 1
COPY_BYTES = 256 << 20
TRAFFIC_FILE = ROOT / "profiles" / "traffic_grid_fused_boole_d5_n65.json"
VALU_FILE = ROOT / "profiles" / "valu_grid_fused_boole_d5_n65.json"
SETTLE_STEPS = 200  # untimed steps before tuning (about 35 ms of GPU work at n=65)


def load_traffic():
    """HBM bytes per launch of the timed kernel from the committed counter summary, or None."""
    try:
        t = json.loads(TRAFFIC_FILE.read_text())
        return t["traffic_bytes_per_launch"]
    except (OSError, ValueError, KeyError):
        return None


def load_valu():
    """Executed VALU work per evaluation of the timed kernel (committed counter pass), or None."""
    try:
        return json.loads(VALU_FILE.read_text())
    except (OSError, ValueError):
        return None


def probe_rates(dev):
    """The device's own rates for the fused grid kernel's loop (tq_probe_rates):
    cycles per wave instruction of FP32 FMA and exp2 chains at the nominal
    clock, and the evaluations/s of the loop body alone."""
    import ctypes

    import torch

    from torchquad_amd._lib import call, stream_ptr

    cus = torch.cuda.get_device_properties(dev).multi_processor_count
    blocks = 32 * cus  # 128 waves per CU
    out = torch.empty(blocks * 256, dtype=torch.float32, device=dev)
    waves_per_simd = blocks * 4 / (4 * cus)
    res = {}
    for kind, iters in ((0, 4000), (1, 1000), (2, 200), (3, 200)):
        for _ in range(2):
            call("tq_probe_rates", kind, blocks, 5, ctypes.c_void_p(out.data_ptr()), stream_ptr(dev))
        torch.cuda.synchronize()
        e0, e1 = _events()
        e0.record()
        call("tq_probe_rates", kind, blocks, iters, ctypes.c_void_p(out.data_ptr()), stream_ptr(dev))
        e1.record()
        torch.cuda.synchronize()
        sec = e0.elapsed_time(e1) / 1e3
        if kind < 2:
            res["fma_cycles" if kind == 0 else "exp2_cycles"] = sec * CLOCK_GHZ * 1e9 / (waves_per_simd * iters * 32)
        else:
            res["loop_evals_per_s" if kind == 3 else "loop_weighted_evals_per_s"] = blocks * 256 * iters * 256 / sec
    return res


def simd_count():
    import torch

    return 4 * torch.cuda.get_device_properties(0).multi_processor_count


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def _events():
    import torch

    return torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)


def bench_mc(c2b, dev, steps):
    """Monte Carlo dim 8, 1e8 samples, FP32 oscillatory (HIP events on the launch stream)."""
    import torch

    import torchquad_amd as tq

    mc = tq.MonteCarlo()
    fo = tq.Oscillatory(c2b["c"], c2b["w"])
    dom8 = torch.tensor(c2b["domain"], dtype=torch.float32, device=dev)
    for _ in range(2):
        mc.integrate(fo, 8, N=c2b["N"], integration_domain=dom8, seed=0)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream(dev)
    e0, e1 = _events()
    k = max(3, steps // 4)
    e0.record(stream)
    for i in range(k):
        mval = mc.integrate(fo, 8, N=c2b["N"], integration_domain=dom8, seed=i)
    e1.record(stream)
    torch.cuda.synchronize()
    mc_ms = e0.elapsed_time(e1) / k
    se = mc.std_error()
    return {
        "evals_per_s": c2b["N"] / (mc_ms / 1e3),
        "ms_per_call": mc_ms,
        "value": float(mval),
        "analytic": c2b["analytic"],
        "abs_err": abs(float(mval) - c2b["analytic"]),
        "std_error": se,
    }


def bench_vegas(c3, dev):
    """VEGAS dim 6, 50 bins/axis, 20 iterations of about 1e6 evals."""
    import torch

    import torchquad_amd as tq

    f = tq.DoubleGaussian(6, width=c3["width"])
    dom = torch.tensor(c3["domain"], dtype=torch.float32, device=dev)
    N = c3["N_increment"] * (c3["iterations"] + 5)
    v = tq.VEGAS()
    # untimed run of the same shape: the integrator keeps its device state
    # between calls of one shape, so the timed calls measure the iterations
    v.integrate(f, 6, N=N, integration_domain=dom, seed=99, N_intervals=c3["N_intervals"],
                max_iterations=c3["iterations"])
    torch.cuda.synchronize()
    times, vals = [], []
    for seed in range(5):
        t0 = time.perf_counter()
        val = float(v.integrate(f, 6, N=N, integration_domain=dom, seed=seed, N_intervals=c3["N_intervals"],
                                max_iterations=c3["iterations"]))
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        vals.append(val)
    t = statistics.median(times)
    return {
        "evals_per_s": v._nr_of_fevals / t,
        "ms_per_integrate": t * 1e3,
        "ms_per_integrate_all": [round(x * 1e3, 3) for x in times],
        "timing": "median of 5 integrates (seeds 0..4) after one untimed integrate of the same shape",
        "evals": int(v._nr_of_fevals),
        "iterations": int(v.it),
        "schedule_note": f"N = {N} = 1e6 x (20 + 5), max_iterations=20, 50 bins: the reference's schedule "
                         "(vegas.py:161-209) raises the evaluations per iteration by N/25 after every five "
                         "stable iterations and stops when the next five would exceed N, here after "
                         f"{int(v.it)} iterations",
        "value": vals[-1],
        "analytic": c3["analytic"],
        "std_error": float(v._get_error()),
    }


def bench_batched(dev, steps):
    """8192 Genz integrals: graph-replayed batch vs a loop of plain calls."""
    import torch

    import torchquad_amd as tq
    from torchquad_amd.autotune import autotune
    from torchquad_amd.batch import Batch
    from torchquad_amd.workloads import genz_batch

    items = genz_batch(8192, seed=0)
    b = Batch(items, dtype=torch.float32, device=dev)
    tile, _ = autotune(b, reps=3, use_cache=False)
    b.run()
    acc = batch_accuracy(items, b.out.double().cpu().numpy())
    g = b.graph()
    for _ in range(3):
        g.replay()
    torch.cuda.synchronize()
    e0, e1 = _events()
    reps = max(10, steps)
    e0.record()
    for _ in range(reps):
        g.replay()
    e1.record()
    torch.cuda.synchronize()
    ms_batch = e0.elapsed_time(e1) / reps

    integrators = {"trapezoid": tq.Trapezoid(), "simpson": tq.Simpson(), "gauss_legendre": tq.GaussLegendre()}
    doms = {d: torch.tensor([[0.0, 1.0]] * d, dtype=torch.float32, device=dev) for d in range(1, 7)}

    def plain_loop(sub):
        for it in sub:
            integrators[it["rule"]].integrate(it["fn"], it["dim"], N=it["N"], integration_domain=doms[it["dim"]])

    plain_loop(items)  # warm pass: per-n Gauss-Legendre tables cached, as the batch plan holds them
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plain_loop(items)
    torch.cuda.synchronize()
    ms_plain = (time.perf_counter() - t0) * 1e3
    return {
        "integrals": len(items),
        "evals": int(b.evals),
        "evals_per_s": b.evals / (ms_batch / 1e3),
        "ms_per_batch": ms_batch,
        "tile_points": tile,
        "plain_loop_ms": ms_plain,
        "plain_loop_evals_per_s": b.evals / (ms_plain / 1e3),
        "speedup_vs_plain_loop": ms_plain / ms_batch,
        **acc,
    }


def batch_accuracy(items, out):
    """FP32 batch results against FP64 separable sums of the same rules (closed
    forms for Gauss-Legendre beyond n = 1000, where the rule is exact to FP64)."""
    import numpy as np

    from oracle import families as fam
    from oracle import grid as og
    from torchquad_amd.batch import RULE_NAMES, n_per_axis

    worst_pos = worst_osc = 0.0
    for it, v in zip(items, out):
        rule, fam_name, d = it["rule"], it["family"], it["dim"]
        n = n_per_axis(RULE_NAMES[rule], d, it["N"])
        a, u, w, dom = it["a"], it["u"], it["w"], it["domain"]
        lo, hi = [0.0] * d, [1.0] * d
        if rule == "gauss_legendre" and n > 1000:
            ref = {"gaussian_peak": lambda: fam.gaussian_peak_integral(a, u, lo, hi),
                   "oscillatory": lambda: fam.oscillatory_integral(a, w, lo, hi),
                   "product_peak": lambda: fam.product_peak_integral(a, u, lo, hi)}[fam_name]()
        elif fam_name == "gaussian_peak":
            ref = og.separable_gaussian_peak(rule, a, u, dom, n)
        elif fam_name == "oscillatory":
            ref = og.separable_oscillatory(rule, a, w, dom, n)
        else:
            ref = og.separable_product(rule, [lambda x, ak=ak, uk=uk: 1.0 / (ak**-2.0 + (x - uk) ** 2)
                                              for ak, uk in zip(a, u)], dom, n)
        if fam_name == "oscillatory":
            worst_osc = max(worst_osc, abs(v - ref))  # |f| <= 1 on a unit-volume box
        else:
            worst_pos = max(worst_pos, abs(v - ref) / abs(ref))
    return {"max_rel_err_positive_families_vs_fp64_oracle": worst_pos,
            "max_abs_err_oscillatory_vs_fp64_oracle": worst_osc,
            "accuracy_note": "FP32 results of the graph-replayed batch against FP64 sums of the same rules; "
                             "oscillatory integrals (|f| <= 1 on [0,1]^d) by absolute error"}


def bench_param_batch(dev, steps):
    """The reference's vectorised-integrand case shape (modular_benchmark.py:945-1011): Simpson,
    dim 1, N=10001 on [0,1], 200 parameter values p = linspace(1, 5, 200); built-in family
    cos(p x) (Oscillatory, w = 0).  Fused parameter batch vs a loop of 200 plain calls vs the
    same object called as an [N, 200] torch integrand (point writer -> fn -> reduce)."""
    import numpy as np
    import torch

    import torchquad_amd as tq

    p = np.linspace(1.0, 5.0, 200)
    fb = tq.Oscillatory(p[:, None], 0.0)
    dom = torch.tensor([[0.0, 1.0]], dtype=torch.float32, device=dev)
    simp = tq.Simpson()

    def timed(fn, reps):
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            r = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / reps, r

    reps = max(20, steps)
    ms_fused, vals = timed(lambda: simp.integrate(fb, 1, N=10001, integration_domain=dom), reps)
    singles = [tq.Oscillatory([q], 0.0) for q in p]
    ms_loop, _ = timed(lambda: [simp.integrate(f1, 1, N=10001, integration_domain=dom) for f1 in singles], 3)
    ms_vec, uvals = timed(lambda: simp.integrate(lambda x: fb(x), 1, N=10001, integration_domain=dom), reps)
    exact = np.sin(p) / p
    return {
        "workload": "Simpson dim=1 N=10001 [0,1] fp32, cos(p x) for 200 p in [1,5]",
        "ms_fused_param_batch": ms_fused,
        "ms_loop_of_200_calls": ms_loop,
        "ms_torch_vectorised_integrand": ms_vec,
        "speedup_fused_vs_loop": ms_loop / ms_fused,
        "max_rel_err_fused_vs_analytic": float(np.max(np.abs(vals.double().cpu().numpy() - exact) / np.abs(exact))),
    }


def bench_c1(c1, dev):
    """Reference's CPU-runnable case on the GPU: Trapezoid and Simpson, d=3, N=101^3, fp64,
    f = exp(x1+x2+x3) on [0,1]^3."""
    import torch

    import torchquad_amd as tq

    f = tq.ExpLinear(3, [1.0, 1.0, 1.0])
    dom = torch.tensor(c1["domain"], dtype=torch.float64, device=dev)
    out = {}
    for name, integ in (("trapezoid", tq.Trapezoid()), ("simpson", tq.Simpson())):
        for _ in range(3):
            integ.integrate(f, 3, N=c1["N"], integration_domain=dom)
        torch.cuda.synchronize()
        e0, e1 = _events()
        reps = 20
        e0.record()
        for _ in range(reps):
            v = integ.integrate(f, 3, N=c1["N"], integration_domain=dom)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        out[name] = {"ms_per_call": ms, "evals_per_s": c1["N"] / (ms / 1e3), "value": float(v),
                     "rel_err_vs_analytic": abs(float(v) - c1["analytic"]) / c1["analytic"]}
    return out


def cpu_ports(cfg, c2b, seconds=3.0):
    """The reference's algorithms on the host for the other configs: torch-CPU
    on every core for C1 and the Monte Carlo config, the numpy VEGAS
    restatement (one thread) for a bounded VEGAS sample."""
    import numpy as np
    import torch

    from oracle import families as fam
    from oracle import torch_cpu
    from oracle import vegas as ov

    out = {}
    cores = os.cpu_count()
    before = torch.get_num_threads()
    torch.set_num_threads(cores)
    c1 = cfg["C1"]
    for rule in ("trapezoid", "simpson"):
        val, _, pts, dt = torch_cpu.grid_rule_slices(rule, torch_cpu.exp_sum_fn(), c1["domain"], 101,
                                                     dtype=torch.float64)
        out[f"c1_{rule}"] = {"evals_per_s": pts / dt, "value": val, "cores": cores,
                             "sample": "full C1 grid 101^3 fp64 (torch-CPU restatement)"}
    n_mc, dt, est = torch_cpu.monte_carlo_oscillatory(c2b["c"], c2b["w"], c2b["domain"], seconds)
    out["mc_d8"] = {"evals_per_s": n_mc / dt, "estimate": est, "cores": cores,
                    "sample": f"{n_mc} samples in chunks of 1e7 (torch.rand + cos, torch-CPU)"}
    torch.set_num_threads(before)
    c3 = cfg["C3"]
    n_v = 2_500_000
    t0 = time.perf_counter()
    r = ov.run(lambda x: fam.double_gaussian(x, c3["width"]), 6, n_v, c3["domain"], seed=0, dtype=np.float32,
               N_intervals=c3["N_intervals"])
    dt = time.perf_counter() - t0
    out["vegas_d6"] = {"evals_per_s": r["fevals"] / dt, "sample": f"VEGAS N={n_v} (numpy restatement)",
                       "cores": 1}
    return out


def bench_serving(dev, rate=2000.0, count=4000):
    """Serving loop: Poisson arrivals of Genz integrals at `rate`/s; latency and throughput."""
    import numpy as np
    import torch

    from torchquad_amd.serve import IntegrationServer, latency_stats, serve_open_loop
    from torchquad_amd.workloads import genz_batch

    items = genz_batch(count, seed=1)

    def req(it):
        return lambda s, arrival: s.submit_grid(it["rule"], it["fn"], it["dim"], N=it["N"], domain=it["domain"],
                                                arrival=arrival)

    server = IntegrationServer(dev, dtype=torch.float32)
    # warm-up: Gauss-Legendre tables, pinned buffers, kernels
    serve_open_loop(server, [req(it) for it in items[:200]], rate=1e6, seed=2)
    server.completed.clear()
    # host-side pauses during the run: garbage-collector passes and the
    # longest scheduling round
    import gc

    gc_pauses = []
    gc_t = [0.0]

    def gc_cb(phase, info):
        if phase == "start":
            gc_t[0] = time.perf_counter()
        else:
            gc_pauses.append((info.get("generation"), time.perf_counter() - gc_t[0]))

    step_max = [0.0]
    orig_step = server.step

    def timed_step():
        t = time.perf_counter()
        r = orig_step()
        step_max[0] = max(step_max[0], time.perf_counter() - t)
        return r

    server.step = timed_step
    # as a serving process does after start-up: long-lived objects leave the
    # collector's generations, so a full pass during serving does not walk
    # the whole heap (a 37 ms pause otherwise set the p99 in this process)
    gc.collect()
    gc.freeze()
    gc.callbacks.append(gc_cb)
    t0 = time.perf_counter()
    tickets = serve_open_loop(server, [req(it) for it in items], rate=rate, seed=3)
    wall = time.perf_counter() - t0
    gc.callbacks.remove(gc_cb)
    server.step = orig_step
    st = latency_stats(tickets)
    gl_code = 3
    st_gl = latency_stats([t for t in tickets if t.rule == gl_code])
    st_nc = latency_stats([t for t in tickets if t.rule != gl_code])
    evals = sum(t.evals for t in tickets)
    # saturation: every request submitted at once, one submit_grid call each
    server.completed.clear()
    t0 = time.perf_counter()
    tickets2 = [req(it)(server, None) for it in items]
    server.run_until_idle(timeout=120)
    wall2 = time.perf_counter() - t0
    # saturation with the same requests as packed arrays of 512 (submit_many)
    from torchquad_amd.workloads import genz_arrays

    arr = genz_arrays(count, seed=1)
    chunk = 512

    def packed():
        ts = [server.submit_many(arr["rule"][k:k + chunk], arr["family"][k:k + chunk], arr["dim"][k:k + chunk],
                                 arr["domain"][k:k + chunk], arr["a"][k:k + chunk], arr["u"][k:k + chunk],
                                 arr["c"][k:k + chunk], N=arr["N"][k:k + chunk]) for k in range(0, count, chunk)]
        server.run_until_idle(timeout=120)
        return ts

    packed()  # warm-up
    server.completed.clear()
    walls = []
    for _ in range(5):
        t0 = time.perf_counter()
        tickets3 = packed()
        walls.append(time.perf_counter() - t0)
    wall3 = statistics.median(walls)
    same = all(np.array_equal(t.result(), np.array([x.result() for x in tickets2[k * chunk:(k + 1) * chunk]]))
               for k, t in enumerate(tickets3))
    # the plain per-call path on the same requests
    import torchquad_amd as tq

    integ = {"trapezoid": tq.Trapezoid(), "simpson": tq.Simpson(), "gauss_legendre": tq.GaussLegendre()}
    doms = {d: torch.tensor([[0.0, 1.0]] * d, dtype=torch.float32, device=dev) for d in range(1, 7)}
    for it in items:  # warm pass (Gauss-Legendre tables)
        integ[it["rule"]].integrate(it["fn"], it["dim"], N=it["N"], integration_domain=doms[it["dim"]])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for it in items:
        integ[it["rule"]].integrate(it["fn"], it["dim"], N=it["N"], integration_domain=doms[it["dim"]])
    torch.cuda.synchronize()
    wall_plain = time.perf_counter() - t0
    return {
        "arrival_rate_per_s": rate,
        "requests": count,
        "p50_ms": st.get("p50_ms"),
        "p99_ms": st.get("p99_ms"),
        "mean_ms": st.get("mean_ms"),
        "max_ms": st.get("max_ms"),
        "gauss_legendre_p50_p99_ms": [st_gl.get("p50_ms"), st_gl.get("p99_ms")],
        "newton_cotes_p50_p99_ms": [st_nc.get("p50_ms"), st_nc.get("p99_ms")],
        "gc_frozen_at_start": True,
        "gc_passes": len(gc_pauses),
        "gc_pause_max_ms": max((p for _, p in gc_pauses), default=0.0) * 1e3,
        "gc_gen2_passes": sum(1 for g, _ in gc_pauses if g == 2),
        "longest_round_ms": step_max[0] * 1e3,
        "evals_per_s_at_rate": evals / wall,
        "batches": server.batches_launched,
        "saturated_single_submit_evals_per_s": sum(t.evals for t in tickets2) / wall2,
        "saturated_single_submit_requests_per_s": count / wall2,
        "saturated_evals_per_s": sum(t.evals for t in tickets3) / wall3,
        "saturated_requests_per_s": count / wall3,
        "saturated_note": f"the {count} requests as packed arrays of {chunk} (submit_many), submitted at once; "
                          "median of 5 drains; values bit-identical to the single submissions: " + str(same),
        "plain_loop_evals_per_s": sum(t.evals for t in tickets2) / wall_plain,
        "saturated_vs_plain_loop": wall_plain / wall3,
        "workload": "Genz integrals (dim 1..6, N 1e3..1e6, trapezoid/simpson/gauss-legendre), numpy PCG64 seed 1",
    }


def bench_grad(c5, dev, steps):
    """1024 Gaussian-peak integrals (GL dim 4, n=32) forward + backward."""
    import numpy as np
    import torch

    from torchquad_amd.autograd import gaussian_peak_integrals
    import torchquad_amd as tq

    rng = np.random.default_rng(c5["seed"])
    B = c5["count"]
    a = rng.uniform(1.0, 5.0, (B, 4))
    u = rng.uniform(-0.5, 0.5, (B, 4))
    lo = rng.uniform(-1.0, 0.0, (B, 4))
    dom = np.stack([lo, lo + 1.0], axis=2)
    ta = torch.tensor(a, dtype=torch.float32, device=dev, requires_grad=True)
    tu = torch.tensor(u, dtype=torch.float32, device=dev, requires_grad=True)
    td = torch.tensor(dom, dtype=torch.float32, device=dev, requires_grad=True)

    def step():
        out = gaussian_peak_integrals(ta, tu, td, n=c5["n"])
        out.sum().backward()
        return out

    for _ in range(3):
        step()
    torch.cuda.synchronize()
    e0, e1 = _events()
    reps = max(10, steps)
    e0.record()
    for _ in range(reps):
        out = step()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    evals = B * c5["n"] ** 4

    # a reference user's loop: one GaussLegendre call per vector, autograd
    # through the points and the torch integrand
    gl = tq.GaussLegendre()
    sub = 16

    def one(i):
        ai = ta[i].detach().clone().requires_grad_(True)
        ui = tu[i].detach().clone().requires_grad_(True)
        di = td[i].detach().clone().requires_grad_(True)
        v = gl.integrate(lambda x: torch.exp(-torch.sum((ai * (x - ui)) ** 2, dim=1)), 4, N=c5["n"] ** 4,
                         integration_domain=di)
        v.backward()

    for i in range(2):
        one(i)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(sub):
        one(i)
    torch.cuda.synchronize()
    ms_loop = (time.perf_counter() - t0) * 1e3 * B / sub

    # the other two Genz families of the batched workload on the same shape
    from torchquad_amd.autograd import oscillatory_integrals, product_peak_integrals

    c_osc = torch.tensor(rng.uniform(0.0, 5.0, (B, 4)), dtype=torch.float32, device=dev, requires_grad=True)
    w_osc = torch.tensor(rng.uniform(0.0, 1.0, B), dtype=torch.float32, device=dev, requires_grad=True)
    families = {}
    for name, fn, p1, p2 in (("oscillatory", oscillatory_integrals, c_osc, w_osc),
                              ("product_peak", product_peak_integrals, ta, tu)):
        def fstep(fn=fn, p1=p1, p2=p2):
            fn(p1, p2, td, n=c5["n"]).sum().backward()

        for _ in range(3):
            fstep()
        torch.cuda.synchronize()
        e0.record()
        for _ in range(reps):
            fstep()
        e1.record()
        torch.cuda.synchronize()
        families[name] = {"ms_forward_backward": e0.elapsed_time(e1) / reps}
    return {
        "vectors": B,
        "evals_per_pass": evals,
        "ms_forward_backward": ms,
        "other_families_same_shape": families,
        "evals_per_s_forward_backward": 2 * evals / (ms / 1e3),
        "first_value": float(out[0]),
        "first_analytic": c5["first_analytic"],
        "per_vector_autograd_loop_ms": ms_loop,
        "per_vector_loop_sample": f"{sub} of {B} vectors timed, scaled to {B}",
        "speedup_vs_autograd_loop": ms_loop / ms,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=65, help="Boole points per axis of the timed step")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mc", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the VEGAS, batched and gradient lines")
    ap.add_argument("--ipw", type=int, default=0,
                    help="items per wave of the timed fused kernel (0: tune before timing; with "
                    "--traffic-probe, the library's automatic choice)")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="run only a 256 MiB elementwise copy (counter calibration) and the timed step, "
                    "for rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE passes")
    ap.add_argument("--profile-section", choices=["mc", "vegas", "batched", "grad", "param", "serving"],
                    help="run only this side section (for rocprofv3 counter passes) and print its result")
    args = ap.parse_args()
    if args.gpus != 1:
        print(
            f"bench.py measures one GPU only (asked for --gpus {args.gpus}); "
            "integrals shard across GPUs without cross-GPU traffic in the hot loop, but only the "
            "one-GPU throughput is measured here.",
            file=sys.stderr,
        )
        sys.exit(2)

    import numpy as np
    import torch

    import torchquad_amd as tq
    from oracle import grid as og

    warnings.simplefilter("ignore")
    cfg = json.loads((ROOT / "tests" / "golden" / "configs.json").read_text())
    c2a, c2b = cfg["C2a"], cfg["C2b"]
    dev = torch.device("cuda", 0)
    torch.cuda.init()
    f = tq.GaussianPeak(c2a["a"], c2a["u"])
    dom = torch.tensor(c2a["domain"], dtype=torch.float32, device=dev)
    boole = tq.Boole()
    n = args.n
    N = n**5
    stream = torch.cuda.current_stream(dev)

    if args.profile_section:
        torch.cuda.init()
        section = {
            "mc": lambda: bench_mc(c2b, dev, args.steps),
            "vegas": lambda: bench_vegas(cfg["C3"], dev),
            "batched": lambda: bench_batched(dev, args.steps),
            "grad": lambda: bench_grad(cfg["C5"], dev, args.steps),
            "param": lambda: bench_param_batch(dev, args.steps),
            "serving": lambda: bench_serving(dev),
        }[args.profile_section]()
        print(json.dumps({"profile_section": args.profile_section, **section}))
        return

    if args.traffic_probe:
        # plain copy of known size: reads and writes COPY_BYTES each
        src = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.neg(src, out=dst)
        torch.cuda.synchronize()

    # Long-lived objects (torch, the package, the configs) leave the cyclic
    # collector's generations, as a serving process does after start-up: a
    # full collection walks torch's heap for tens of milliseconds, and one
    # landing inside a timed loop stalls the launching thread while the HIP
    # events keep counting (an 8 ms "call" for a 0.46 ms Monte Carlo kernel).
    import gc

    gc.collect()
    gc.freeze()

    # The GPU raises its clocks over the first tens of milliseconds of
    # sustained work; run the step untimed until then, so that tuning and the
    # timed steps both see the sustained clock.
    for _ in range(SETTLE_STEPS):
        boole.integrate(f, 5, N=N, integration_domain=dom)
    torch.cuda.synchronize()

    # project autotuner: items per wave of the fused kernel for this geometry,
    # chosen before the timed region
    from torchquad_amd import _lib
    from torchquad_amd.autotune import tune_grid

    if args.ipw > 0:
        from torchquad_amd.autotune import set_grid_ipw

        set_grid_ipw(_lib.RULE_BOOLE, f.descriptor(5).family, 5, n, dom.dtype, args.ipw)
    # counter passes (--traffic-probe) keep the library's automatic choice
    # unless --ipw fixes one, so that every pass sees the same launch
    ipw = args.ipw if args.traffic_probe else tune_grid(_lib.RULE_BOOLE, f, 5, n, dom)[0]

    for _ in range(args.warmup):
        boole.integrate(f, 5, N=N, integration_domain=dom)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.steps):
        val = boole.integrate(f, 5, N=N, integration_domain=dom)
    e1.record(stream)
    torch.cuda.synchronize()
    ms_total = e0.elapsed_time(e1)
    ms_per_step = ms_total / args.steps
    evals_per_s = N / (ms_per_step / 1e3)
    val = float(val)
    if args.traffic_probe:
        print(json.dumps({"traffic_probe": True, "ms_per_step": ms_per_step, "copy_bytes": COPY_BYTES}))
        return
    rel_err = abs(val - c2a["analytic"]) / abs(c2a["analytic"])
    oracle_same_rule = og.separable_gaussian_peak("boole", c2a["a"], c2a["u"], c2a["domain"], n)
    rel_vs_oracle = abs(val - oracle_same_rule) / abs(oracle_same_rule)

    # per-call latency of the literal 17^5 configuration (host call -> result on device)
    ipw17, _ = tune_grid(_lib.RULE_BOOLE, f, 5, 17, dom)
    lat = []
    N17 = 17**5
    for _ in range(5):
        boole.integrate(f, 5, N=N17, integration_domain=dom)
    torch.cuda.synchronize()
    for _ in range(200):
        t0 = time.perf_counter()
        r17 = boole.integrate(f, 5, N=N17, integration_domain=dom)
        torch.cuda.synchronize()
        lat.append((time.perf_counter() - t0) * 1e3)
    lat.sort()
    p50 = lat[len(lat) // 2]
    p99 = lat[int(len(lat) * 0.99) - 1]

    # Monte Carlo, dim 8, 1e8 samples
    mc_info = None if args.no_mc else bench_mc(c2b, dev, args.steps)

    extra = {}
    if not args.no_extra:
        extra["vegas_d6"] = bench_vegas(cfg["C3"], dev)
        extra["batched_c4"] = bench_batched(dev, args.steps)
        extra["param_batch_simpson_1d"] = bench_param_batch(dev, args.steps)
        extra["grad_c5"] = bench_grad(cfg["C5"], dev, args.steps)
        extra["serving_c4"] = bench_serving(dev)
        extra["c1_fp64_d3_n101"] = bench_c1(cfg["C1"], dev)
        if not args.no_cpu:
            extra["cpu_port_other_configs"] = cpu_ports(cfg, c2b)

    # Roofline of the timed kernel.  One step is one launch of the fused kernel
    # (the last workgroup finishes the sum), so the step time measured above
    # with HIP events on the launch stream is the kernel's duration per launch.
    # Executed FP32 operations per evaluation come from the committed counter
    # pass (SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F32 x 64 lanes, FMA = 2); the
    # VALU issue fraction counts every wave instruction as 4 cycles of its
    # SIMD and a transcendental as 8.
    valu = load_valu()
    kernel_s = ms_per_step / 1e3
    alg_tflops = FLOPS_PER_EVAL_GRID_GAUSSIAN * evals_per_s / 1e12
    roofline = {
        "bound": "valu",
        "kernel": "k_grid_fused<float, GAUSSIAN_PEAK, table, rowtab, 8 waves> (one launch per step)",
        "achieved": None,
        "peak": PEAK_FP32_TFLOPS,
        "unit": "TFLOP/s",
        "frac": None,
        "bound_note": "vector ALU (no HBM stream or matrix product on this path): MI355X's FP32 157.3 TFLOP/s; "
                      "each evaluation is one add, one exp2 and one FMA, and the exp2 issues at about a quarter "
                      "of the FMA rate (see attainable), so the kernel is bound by its exp2 issue well below "
                      "the FMA peak",
        "traffic": load_traffic(),
        "algorithmic_equivalent": {
            "flops_per_eval": FLOPS_PER_EVAL_GRID_GAUSSIAN,
            "tflops": alg_tflops,
            "note": "4d+3 arithmetic + 1 exp per evaluation at d=5 as if each point were evaluated from "
                    "scratch; not a hardware rate (the kernel shares row and lane partial sums)",
        },
    }
    rates = probe_rates(dev)
    roofline["attainable"] = {
        "loop_evals_per_s": rates["loop_evals_per_s"],
        "frac_of_loop": evals_per_s / rates["loop_evals_per_s"],
        "fma_cycles_per_wave_instruction": rates["fma_cycles"],
        "exp2_cycles_per_wave_instruction": rates["exp2_cycles"],
        "loop_weighted_evals_per_s": rates["loop_weighted_evals_per_s"],
        "note": "tq_probe_rates on this GPU, cycles at the nominal 2.4 GHz: independent FP32 FMA and exp2 "
                "chains, and the grid loop's body alone as the kernel runs it for this family (state + "
                "4-byte LDS table term with the weight folded in, exp2, add, 16 points per trip) with none "
                "of the kernel's tables, rows, flushes or tail; loop_weighted_evals_per_s: the same body "
                "with (term, weight) entries and a weighted FMA, the form of the non-folding families",
    }
    if valu:
        flops_eval = valu["executed_fp32_flops_per_eval"]
        achieved = flops_eval * evals_per_s / 1e12
        trans = valu["trans_f32_insts_per_eval"]
        other = valu["valu_insts_per_eval"] - trans
        # issue time of the executed instructions at the rates measured on this
        # GPU (every non-transcendental instruction at the FMA rate)
        issue_s = (other * rates["fma_cycles"] + trans * rates["exp2_cycles"]) * N / (
            simd_count() * CLOCK_GHZ * 1e9)
        roofline.update({
            "achieved": achieved,
            "frac": achieved / PEAK_FP32_TFLOPS,
            "executed_fp32_flops_per_eval": flops_eval,
            "valu_issue_frac": issue_s / kernel_s,
            "valu_note": "executed work per evaluation from " + VALU_FILE.name + " (" + valu["command"] +
                         "); valu_issue_frac = (non-transcendental VALU wave instructions x the measured FMA "
                         "cycles + transcendentals x the measured exp2 cycles) / (SIMDs x 2.4 GHz x measured "
                         "time per launch)",
        })

    cpu = None
    if not args.no_cpu:
        # the reference's algorithm on torch CPU (linspace axes, ij meshgrid
        # points, the integrand, the per-axis Boole collapse) for the same
        # workload, FP32, on every host core; bounded to about cpu_seconds by
        # running whole slices of the outermost axis
        from oracle import torch_cpu

        cores = os.cpu_count()
        threads_before = torch.get_num_threads()
        torch.set_num_threads(cores)
        fn_cpu = torch_cpu.gaussian_peak_fn(c2a["a"], c2a["u"])
        torch_cpu.grid_rule_slices("boole", fn_cpu, c2a["domain"], 17)  # warm-up
        cval, slices, pts, dt = torch_cpu.grid_rule_slices("boole", fn_cpu, c2a["domain"], n,
                                                           seconds=args.cpu_seconds)
        torch.set_num_threads(threads_before)
        cpu = {
            "value": pts / dt,
            "unit": "integrand evals/s",
            "cores": cores,
            "kind": "port",
            "sample": f"torch-CPU restatement of the reference path (oracle/torch_cpu.py): Boole d=5 n={n} "
            f"FP32 Gaussian peak, {slices} of {n} outer-axis slices ({pts} evals, {dt:.1f} s), "
            f"torch.set_num_threads({cores}), {cpu_model()}"
            + (f", full-grid value {cval:.9g}" if cval is not None else ""),
        }

    out = {
        "metric": "integrand evals/s/GPU at 1e-6 rel err, plain vs batched; p50/p99 latency",
        "value": evals_per_s,
        "unit": "integrand evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Genz Gaussian-peak parameters, numpy PCG64 seed 0)",
        "config": {
            "workload": f"plain Boole dim=5 [-1,1]^5 fp32 Gaussian peak, n={n}/axis ({N} evals per step)",
            "items_per_wave": ipw,
            "items_per_wave_n17": ipw17,
            "rel_err_vs_analytic": rel_err,
            "rel_err_vs_fp64_oracle_same_rule": rel_vs_oracle,
            "n17_latency_ms_p50": p50,
            "n17_latency_ms_p99": p99,
            "n17_value": float(r17),
            "monte_carlo_d8_1e8": mc_info,
            **extra,
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
