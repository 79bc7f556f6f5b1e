"""Benchmark of the integration hot path on one MI355X.

Workload (the configuration the headline metric is quoted on): plain
single-integral GPU path, Boole rule, dim=5 on [-1,1]^5, fp32, Gaussian peak
exp(-sum a_i^2 (x_i-u_i)^2) with a_i~U(1,5), u_i~U(-0.5,0.5) (numpy PCG64,
seed 0).  The metric is "integrand evals/s at 1e-6 relative error", so the
timed step is one Boole().integrate call at n=65 points per axis (65^5 =
1.16e9 evals), the smallest Boole-valid n whose result is within 1e-6 of the
analytic value for these parameters; n=17 (the literal 17^5 grid) is
reported as per-call latency p50/p99.  Monte Carlo dim=8 with 1e8 samples
(oscillatory integrand) is reported beside it.

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
# FLOPs per evaluation executed by the fused grid kernel for the Gaussian peak
# (loop-axis fold: 2 FMA; exp: 1 mul + 1 exp; weighted accumulate: 1 FMA).
FLOPS_PER_EVAL_GRID_GAUSSIAN = 8


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=65, help="Boole points per axis of the timed step")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mc", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(
            f"bench.py measures one GPU only (asked for --gpus {args.gpus}); "
            "multi-GPU sharding is described in DESIGN.md, not run.",
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
    rel_err = abs(val - c2a["analytic"]) / abs(c2a["analytic"])
    oracle_same_rule = og.separable_gaussian_peak("boole", c2a["a"], c2a["u"], c2a["domain"], n)
    rel_vs_oracle = abs(val - oracle_same_rule) / abs(oracle_same_rule)

    # per-call latency of the literal 17^5 configuration (host call -> result on device)
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
    mc_info = None
    if not args.no_mc:
        mc = tq.MonteCarlo()
        fo = tq.Oscillatory(c2b["c"], c2b["w"])
        dom8 = torch.tensor(c2b["domain"], dtype=torch.float32, device=dev)
        for _ in range(2):
            mc.integrate(fo, 8, N=c2b["N"], integration_domain=dom8, seed=0)
        torch.cuda.synchronize()
        k = max(3, args.steps // 4)
        e0.record(stream)
        for i in range(k):
            mval = mc.integrate(fo, 8, N=c2b["N"], integration_domain=dom8, seed=i)
        e1.record(stream)
        torch.cuda.synchronize()
        mc_ms = e0.elapsed_time(e1) / k
        se = mc.std_error()
        mc_info = {
            "evals_per_s": c2b["N"] / (mc_ms / 1e3),
            "ms_per_call": mc_ms,
            "value": float(mval),
            "analytic": c2b["analytic"],
            "abs_err": abs(float(mval) - c2b["analytic"]),
            "std_error": se,
        }

    achieved_tflops = FLOPS_PER_EVAL_GRID_GAUSSIAN * evals_per_s / 1e12
    roofline = {
        "bound": "valu",
        "kernel": "k_grid_fused<float, GAUSSIAN_PEAK, table>",
        "achieved": achieved_tflops,
        "peak": PEAK_FP32_TFLOPS,
        "unit": "TFLOP/s",
        "frac": achieved_tflops / PEAK_FP32_TFLOPS,
        "flops_per_eval": FLOPS_PER_EVAL_GRID_GAUSSIAN,
        "traffic": None,
    }

    cpu = None
    if not args.no_cpu:
        # oracle restatement (materialised grid + per-axis collapse, numpy FP64)
        # of the same rule/integrand on the 17^5 grid, repeated for ~cpu_seconds
        from oracle import families as fam

        a = np.array(c2a["a"])
        u = np.array(c2a["u"])
        calls = 0
        t0 = time.perf_counter()
        while True:
            og.integrate("boole", lambda x: fam.gaussian_peak(x, a, u), 5, N17, c2a["domain"])
            calls += 1
            if time.perf_counter() - t0 >= args.cpu_seconds:
                break
        dt = time.perf_counter() - t0
        cpu = {
            "value": calls * N17 / dt,
            "unit": "integrand evals/s",
            "cores": 1,
            "kind": "port",
            "sample": f"oracle.grid.integrate('boole', GaussianPeak d=5, N=17^5) x{calls} "
            f"({dt:.1f} s, numpy FP64, {cpu_model()}, os.cpu_count()={os.cpu_count()})",
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
            "rel_err_vs_analytic": rel_err,
            "rel_err_vs_fp64_oracle_same_rule": rel_vs_oracle,
            "n17_latency_ms_p50": p50,
            "n17_latency_ms_p99": p99,
            "n17_value": float(r17),
            "monte_carlo_d8_1e8": mc_info,
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
