"""Writes tests/golden/ipm_single_lp.json: the oracle's result on the single-LP
workload (m=2,000, n=5,000, 5 nonzeros per column, seeds 1-2), with the
independent HiGHS optimum beside it.  Minutes of CPU; run from the repo root:
    python tests/golden/make_single_lp_golden.py
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "indigenous-cpu-gpu-optimization-engine-for-indian-industry_amd"))
import lpgen  # noqa: E402
import oracle as O  # noqa: E402


def main():
    out = {}
    for seed in (1, 2):
        lp = lpgen.random_lp(seed, 2000, 5000)
        t = time.time()
        r = O.ipm_solve(lp)
        dt = time.time() - t
        rec = {"status": r.status, "iterations": r.iterations, "objective": r.objective,
               "rp_norm": r.rp_norm, "rd_norm": r.rd_norm, "mu": r.mu, "oracle_seconds": dt}
        try:
            import scipy.sparse as sp
            from scipy.optimize import linprog
            A = sp.csr_matrix((lp.val, lp.col, lp.row_ptr), shape=(lp.m, lp.n))
            h = linprog(lp.obj, A_eq=A, b_eq=lp.rhs, bounds=[(0, None)] * lp.n, method="highs")
            rec["highs_objective"] = float(h.fun)
        except Exception as e:  # scipy missing: the oracle value stands alone
            rec["highs_error"] = str(e)
        out[str(seed)] = rec
        print(seed, rec, flush=True)
    with open(os.path.join(HERE, "ipm_single_lp.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
