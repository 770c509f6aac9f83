"""Extract small label/number fixtures from the reference repository.

Run once in the build container (where the reference is mounted read-only):
    python tests/golden/make_fixtures.py /path/to/reference
It reads only data files (CSV / JSON / a README table); no reference code is
imported or run.  ICs are NOT stored: they regenerate bit-exactly from
torch.manual_seed(42); torch.rand(n, S) * (max - min) + min (float32, CPU),
which this script asserts against the CSV columns before dropping them.
"""

import csv
import json
import os
import sys

import numpy as np
import torch

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))
TI = os.path.join(REF, "tests", "integration")


def read_csv(path):
    with open(path) as f:
        rows = list(csv.DictReader(f))
    return rows


def regen(n, lo, hi, seed=42):
    torch.manual_seed(seed)
    lo_t = torch.tensor(lo, dtype=torch.float32)
    hi_t = torch.tensor(hi, dtype=torch.float32)
    return (torch.rand(n, len(lo)) * (hi_t - lo_t) + lo_t).numpy()


def check_ics(rows, lo, hi):
    cols = [k for k in rows[0] if k.startswith("x")]
    ics = np.array([[float(r[c]) for c in cols] for r in rows], dtype=np.float32)
    want = regen(len(rows), lo, hi)
    assert np.array_equal(ics, want), f"IC regeneration mismatch ({np.abs(ics - want).max()})"


cases = {}
pi = np.pi
asin = np.arcsin(0.5)
specs = {
    "pendulum_case1": (os.path.join(TI, "pendulum/ground_truths/case1/main_pendulum_case1.csv"),
                       [-pi + asin, -10.0], [pi + asin, 10.0],
                       os.path.join(TI, "pendulum/main_pendulum_case1.json")),
    "duffing_main": (os.path.join(TI, "duffing/ground_truths/main/main_duffing.csv"),
                     [-1.0, -0.5], [1.0, 1.0],
                     os.path.join(TI, "duffing/main_duffing_supervised.json")),
    "lorenz_main": (os.path.join(TI, "lorenz/ground_truths/main/main_lorenz.csv"),
                    [-10.0, -20.0, 0.0], [10.0, 20.0, 0.0],
                    os.path.join(TI, "lorenz/main_lorenz.json")),
}
for name, (path, lo, hi, bs_json) in specs.items():
    rows = read_csv(path)
    check_ics(rows, lo, hi)
    with open(bs_json) as f:
        bs = json.load(f)
    cases[name] = {
        "source": os.path.relpath(path, REF),
        "bs_source": os.path.relpath(bs_json, REF),
        "n": len(rows), "seed": 42, "min_limits": lo, "max_limits": hi,
        "labels": [r["label"] for r in rows],
        "bs": bs["data"], "label_map": bs.get("label_map", {}),
    }

# pendulum T sweep (case 2): 20 x 2000, labels only
idx = read_csv(os.path.join(TI, "pendulum/ground_truths/case2/parameter_index.csv"))
sweep = []
for r in idx:
    rows = read_csv(os.path.join(TI, "pendulum/ground_truths/case2", r["filename"]))
    sweep.append({"T": float(r["parameter_value"]), "labels": [x["label"] for x in rows]})
first = read_csv(os.path.join(TI, "pendulum/ground_truths/case2/param_001.csv"))
check_ics(first, [-pi + asin, -10.0], [pi + asin, 10.0])
cases["pendulum_case2_T_sweep"] = {
    "source": "tests/integration/pendulum/ground_truths/case2",
    "n": len(first), "seed": 42, "min_limits": [-pi + asin, -10.0], "max_limits": [pi + asin, 10.0],
    "note": "ICs of every T share the seed-42 draw over the T=0.5 region (asserted for param_001)",
    "sweep": sweep,
}

with open(os.path.join(OUT, "label_fixtures.json"), "w") as f:
    json.dump(cases, f)

# zigode README tables (src/zigode/README.md:59-99), float64, 5 ICs,
# t in [0, 1000], 1000 save points, rtol 1e-8, atol 1e-6.
zig = {
    "source": "src/zigode/README.md:59-99; tolerances src/zigode/test_zig_solver.py:36-41",
    "ics": [[0.4, 0.0], [2.7, 0.0], [0.5, -5.0], [-2.0, 3.0], [1.0, 0.0]],
    "params": {"alpha": 0.1, "T": 0.5, "K": 1.0},
    "t_span": [0.0, 1000.0], "t_steps": 1000, "rtol": 1e-8, "atol": 1e-6,
    "final_theta": {
        "zig": [0.523617, 4947.699671, 4904.608349, 4975.221288, 0.523583],
        "torchdiffeq": [0.523599, 4947.703034, 4904.614031, 4975.223032, 0.523599],
        "jax": [0.523614, 4947.728516, 4904.652344, 4975.236816, 0.523579],
    },
    "max_abs_err_zig_vs_torchdiffeq": [2.03e-05, 3.61e-03, 6.10e-03, 1.77e-03, 1.74e-05],
    "early_t10_err_zig_vs_torchdiffeq": [2.97e-06, 1.42e-05, 6.89e-07, 1.51e-05, 2.26e-06],
    "tolerances": {"CROSS_SOLVER_ATOL": 0.05, "FIXED_POINT_ATOL": 1e-3,
                   "LIMIT_CYCLE_OMEGA_ATOL": 1e-2, "EARLY_TIME_ATOL": 1e-3},
}
with open(os.path.join(OUT, "zigode_pendulum_golden.json"), "w") as f:
    json.dump(zig, f, indent=1)

# Rossler network: paper S_B (setup_rossler_network_system.py:25-31) and the
# reference's own n=500 runs (artifacts/results/rossler_network_k_sweep_comparison.json)
ross = {"source": "case_studies/rossler_network/setup_rossler_network_system.py:25-31",
        "K": [0.119, 0.139, 0.159, 0.179, 0.198, 0.218, 0.238, 0.258, 0.278, 0.297, 0.317],
        "S_B_paper": [0.226, 0.274, 0.330, 0.346, 0.472, 0.496, 0.594, 0.628, 0.656, 0.694, 0.690]}
p = os.path.join(REF, "artifacts/results/rossler_network_k_sweep_comparison.json")
if os.path.exists(p):
    with open(p) as f:
        ross["reference_run_n500"] = json.load(f)
with open(os.path.join(OUT, "rossler_golden.json"), "w") as f:
    json.dump(ross, f, indent=1)
print("fixtures written")
