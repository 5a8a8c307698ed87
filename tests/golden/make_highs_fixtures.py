"""Writes highs_objectives.json: optimal objectives of seeded generator LPs,
computed by scipy's HiGHS (an independent LP code) in this container, for
the GPU tests on the box (where solving them again would hold the GPU).

Cases:
  * batch instances of the 4,096-LP workload (seed 1000+i) spread over the
    whole range of m, the largest ones near m = 2,000;
  * the large power-law LP's generator at reduced sizes (m x 2.5m, column
    counts on [2, 200], seed 3);
  * the first 12 held-out 200 x 300 transportation LPs (seeds 1,000,000+),
    and ten more held-out ones whose iterates stall near a relative gap
    of 1e-7 with the support of the optimal face still ambiguous (a basic
    column with x_j < s_j, or a weakly basic one carrying a negative
    reduced cost after the projection).
Run:  python tests/golden/make_highs_fixtures.py
"""
import json
import os
import sys
import time

import numpy as np
import scipy.sparse as sp
from scipy.optimize import linprog

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "indigenous-cpu-gpu-optimization-engine-for-indian-industry_amd"))
import lpgen  # noqa: E402

OUT = os.path.join(HERE, "highs_objectives.json")


def highs(lp, method):
    A = sp.csr_matrix((lp.val, lp.col, lp.row_ptr), shape=(lp.m, lp.n))
    t = time.time()
    r = linprog(lp.obj, A_eq=A, b_eq=lp.rhs, bounds=(0, None), method=method,
                options={"primal_feasibility_tolerance": 1e-10, "dual_feasibility_tolerance": 1e-10})
    assert r.status == 0, r.message
    return float(r.fun), time.time() - t


def batch_ids():
    ms = np.array([lpgen.batch_lp(i).m for i in range(4096)])
    order = np.argsort(ms)
    # 4 largest, and 8 spread over the quantiles
    pick = list(order[-4:]) + [order[int(q * 4095)] for q in (0.0, 0.15, 0.3, 0.45, 0.6, 0.75, 0.85, 0.95)]
    return sorted(set(int(i) for i in pick)), ms


def main():
    data = json.load(open(OUT)) if os.path.exists(OUT) else {}
    ids, ms = batch_ids()
    for i in ids:
        key = f"batch_{i}"
        if key in data:
            continue
        lp = lpgen.batch_lp(i)
        obj, dt = highs(lp, "highs")
        data[key] = {"m": int(lp.m), "n": int(lp.n), "objective": obj, "seconds": round(dt, 2)}
        print(key, data[key], flush=True)
        json.dump(data, open(OUT, "w"), indent=1)
    for m in (2000, 5000):
        key = f"powerlaw_{m}x{int(2.5 * m)}"
        if key in data:
            continue
        lp = lpgen.powerlaw_lp(3, m, int(2.5 * m), 2.179, 2, 200)
        obj, dt = highs(lp, "highs-ipm")
        data[key] = {"m": m, "n": int(2.5 * m), "seed": 3, "cmax": 200, "objective": obj, "seconds": round(dt, 2)}
        print(key, data[key], flush=True)
        json.dump(data, open(OUT, "w"), indent=1)


TRANSPORT_HARD = [1_000_101, 1_000_136, 1_000_150, 1_000_174, 1_000_179, 1_000_200, 1_000_207, 1_000_813, 1_000_963,
                  1_000_984]


def transport_main():
    """HiGHS optima of the first held-out transportation LPs (200 x 300)."""
    data = json.load(open(OUT))
    for seed in list(range(1_000_000, 1_000_012)) + TRANSPORT_HARD:
        key = f"transport_{seed}"
        if key in data:
            continue
        lp = lpgen.transport_lp(seed)
        A = sp.csr_matrix((lp.val, lp.col, lp.row_ptr), shape=(lp.m, lp.n))
        le = lp.sense == lpgen.LE
        t = time.time()
        r = linprog(lp.obj, A_ub=A[le], b_ub=lp.rhs[le], A_eq=A[~le], b_eq=lp.rhs[~le], bounds=(0, None),
                    method="highs", options={"primal_feasibility_tolerance": 1e-10,
                                             "dual_feasibility_tolerance": 1e-10})
        assert r.status == 0, r.message
        data[key] = {"m": int(lp.m), "n": int(lp.n), "objective": float(r.fun), "seconds": round(time.time() - t, 2)}
        print(key, data[key], flush=True)
        json.dump(data, open(OUT, "w"), indent=1)


if __name__ == "__main__":
    main()
    transport_main()
