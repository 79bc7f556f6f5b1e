"""Writes the golden fixtures of the test-suite.

reference_vectors.json: known answers held by the reference's own tests and
docs (values transcribed as data, with their file:line), and the Random123
known-answer vectors of Philox-4x32-10.
configs.json: the seeded synthetic inputs of the benchmark configurations
(numpy PCG64, seed 0) and their closed-form integrals.

Run from the repository root: python tests/golden/make_golden.py
"""
import json
import math
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))

from oracle import families as fam  # noqa: E402


def reference_vectors():
    return {
        "philox4x32_10_kat": {
            "source": "Random123 kat_vectors (Salmon et al., SC'11)",
            "cases": [
                {"ctr": [0, 0, 0, 0], "key": [0, 0], "out": [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]},
                {"ctr": [0xFFFFFFFF] * 4, "key": [0xFFFFFFFF] * 2,
                 "out": [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]},
                {"ctr": [0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], "key": [0xA4093822, 0x299F31D0],
                 "out": [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]},
            ],
        },
        "vegas_map_interval": {
            "source": "tests/vegas_map_test.py:24-48",
            "N_intervals": 20,
            "y": [[0.8121, 0.4319, 0.1612], [0.4746, 0.6501, 0.9241], [0.6143, 0.0724, 0.5818]],
            "ID": [[16, 8, 3], [9, 13, 18], [12, 1, 11]],
            "offset": [[0.2420, 0.6380, 0.2240], [0.4920, 0.0020, 0.4820], [0.2860, 0.4480, 0.6360]],
            "offset_tol": 6e-5,
        },
        "vegas_smooth_map": {
            "source": "tests/vegas_map_test.py:84-100",
            "weights": [[0.0, 0.0, 0.0, 1.0, 1.0, 1.0], [0.0, 0.0, 0.0, 1.0, 0.0, 0.0]],
            "counts": [[1] * 6, [1] * 6],
            "alpha": 0.5,
            "expected": [[0.0, 0.0, 0.54913316, 0.75820765, 0.77899047, 0.77899047],
                         [-0.0, -0.0, 0.64868024, 0.93220967, 0.64868024, -0.0]],
            "tol": 3e-7,
        },
        "vegas_map_counts": {
            "source": "tests/vegas_map_test.py:55-81",
            "dim": 3, "N_intervals": 20, "N_per_dim": 100, "y_max": 0.99999, "count": 50000,
        },
        "vegas_strat_fresh": {
            "source": "tests/vegas_stratification_test.py:157-172",
            "N_increment": 1000, "dim": 3, "N_strat": 6, "nevals_exp": 4000, "nh": 18,
        },
        "add_at_indices": {
            "source": "tests/utils_integration_test.py:109-115",
            "target": [0, 0, 0], "indices": [2, 1, 1, 2], "source_values": [1, 10, 100, 1000],
            "expected": [0, 110, 1001],
        },
        "tutorial_outputs": {
            "source": "docs/source/tutorial.rst (fp32 GPU outputs)",
            "cases": [
                {"rule": "trapezoid", "integrand": "exp(x)*x^2", "dim": 1, "N": 101, "domain": [[0, 2]],
                 "value": 12.780082702636719, "cite": "tutorial.rst:222"},
                {"rule": "simpson", "integrand": "exp(x)*x^2", "dim": 1, "N": 101, "domain": [[0, 2]],
                 "value": 12.778112411499023, "cite": "tutorial.rst:238"},
                {"rule": "trapezoid", "integrand": "sum sin(x)", "dim": 10, "N": 59049, "domain": [[0, 1]] * 10,
                 "value": 4.500804901123047, "cite": "tutorial.rst:318"},
                {"rule": "simpson", "integrand": "sum sin(x)", "dim": 10, "N": 59049, "domain": [[0, 1]] * 10,
                 "value": 4.598623752593994, "cite": "tutorial.rst:331"},
                {"rule": "boole", "integrand": "sum sin(x)", "dim": 10, "N": 59049, "domain": [[0, 1]] * 10,
                 "value": 4.596974849700928, "cite": "tutorial.rst:344"},
            ],
            "rel_tol_fp32": 1e-6,
        },
        "gradients": {
            "source": "tests/gradient_test.py:187-259",
            "v_function_domain": {"domain": [[-1.0, 1.0]], "integral": 2.0, "grad": [[-2.0, 2.0]],
                                  "tol_integral": 1e-2, "tol_grad": 2e-2},
            "poly_domain": {"domain": [[0.0, 1.0], [0.0, 2.0]], "integral": 12.0,
                            "grad": [[-10.0, 14.0], [-2.0, 14.0]], "tol_integral": 8e-2, "tol_grad": 0.1},
            "poly_coeffs": {"domain": [[0.0, 1.0], [0.0, 2.0]], "param": [1.0, 2.0, 3.0], "integral": 12.0,
                            "grad": [2.0, 1.0, 8.0 / 3.0], "tol_integral": 8e-2, "tol_grad": 5e-2},
            "v_offset": {"domain": [[-5.0, 3.0]], "param": 2.0, "integral": 34.0, "grad": 4.0,
                         "tol_integral": 0.2, "tol_grad": 0.15},
            "Ns_1d": [149, 149, 149, 99997, 99997],
            "Ns_2d": [549, 121, 81, 99997, 99997],
        },
        "N_adjust": {
            "source": "tests/simpson_test.py:53, tests/boole_test.py:36, integration_grid_test.py:48-95",
            "cases": [{"rule": "simpson", "dim": 3, "N": 1076890, "n": 101},
                      {"rule": "boole", "dim": 3, "N": 1076890, "n": 101},
                      {"rule": "trapezoid", "dim": 3, "N": 64, "n": 4}],
        },
        "vegas_special": {
            "source": "tests/vegas_test.py:164-184",
            "zero_integral": 0.0, "constant_value": 10.0, "constant_domain": [[0.0, 3.0]] * 2,
            "constant_integral": 90.0, "constant_tol": 1e-13,
        },
    }


def configs(seed=0):
    rng = np.random.default_rng(seed)
    # C2a: Boole d=5, [-1,1]^5, Gaussian peak a~U(1,5), u~U(-0.5,0.5)
    a = rng.uniform(1.0, 5.0, 5)
    u = rng.uniform(-0.5, 0.5, 5)
    c2a = {"rule": "boole", "dim": 5, "n": 17, "domain": [[-1.0, 1.0]] * 5, "a": a.tolist(), "u": u.tolist(),
           "analytic": fam.gaussian_peak_integral(a, u, [-1.0] * 5, [1.0] * 5)}
    # C2b: MC d=8, N=1e8, oscillatory cos(2 pi w + sum c x), c~U(0,2), w~U(0,1), domain [0,1]^8
    c = rng.uniform(0.0, 2.0, 8)
    w = float(rng.uniform(0.0, 1.0))
    c2b = {"dim": 8, "N": 100_000_000, "domain": [[0.0, 1.0]] * 8, "c": c.tolist(), "w": w,
           "analytic": fam.oscillatory_integral(c, w, [0.0] * 8, [1.0] * 8)}
    # C1: Trapezoid / Simpson d=3, [0,1]^3, N=101^3, f=exp(x1+x2+x3)
    c1 = {"dim": 3, "N": 101**3, "domain": [[0.0, 1.0]] * 3, "analytic": (math.e - 1.0) ** 3}
    # C3: VEGAS d=6, [0,1]^6 double Gaussian width 0.1 centres 1/3, 2/3
    c3 = {"dim": 6, "domain": [[0.0, 1.0]] * 6, "N_intervals": 50, "N_increment": 1_000_000,
          "iterations": 20, "width": 0.1,
          "analytic": fam.double_gaussian_integral(6, [0.0] * 6, [1.0] * 6)}
    # C5: GL d=4 n=32, 1024 parameter vectors, domain [l, l+1]
    rng5 = np.random.default_rng(seed + 5)
    a5 = rng5.uniform(1.0, 5.0, (1024, 4))
    u5 = rng5.uniform(-0.5, 0.5, (1024, 4))
    l5 = rng5.uniform(-1.0, 0.0, (1024, 4))
    c5 = {"dim": 4, "n": 32, "count": 1024, "seed": seed + 5,
          "first_analytic": fam.gaussian_peak_integral(a5[0], u5[0], l5[0], l5[0] + 1.0)}
    return {"seed": seed, "C1": c1, "C2a": c2a, "C2b": c2b, "C3": c3, "C5": c5}


if __name__ == "__main__":
    (HERE / "reference_vectors.json").write_text(json.dumps(reference_vectors(), indent=1))
    (HERE / "configs.json").write_text(json.dumps(configs(), indent=1))
    print("wrote", HERE / "reference_vectors.json", HERE / "configs.json")
