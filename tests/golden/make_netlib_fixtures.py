"""Writes the Netlib fixtures: the reference's 21 Netlib MPS models
(benchmarks/netlib/*.mps, public benchmark data) gzipped under
tests/golden/netlib/, and tests/golden/netlib.json with, per model, the
dimensions the reference's own parser read (benchmarks/results/
netlib-baseline-p0.csv: vars, cons, nnz), the reference's own solver result
(status, objective, engine) from that file, the published optimum of the
reference's table (benchmarks/netlib_reference.cpp:80-105), and the optimum
of the parsed LP by HiGHS (scipy.optimize.linprog), an independent solver.

Run in the container that has /root/reference:
    python tests/golden/make_netlib_fixtures.py
"""
import csv
import gzip
import json
import math
import os
import re
import shutil
import sys

import numpy as np
import scipy.optimize as so
import scipy.sparse as sp

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "indigenous-cpu-gpu-optimization-engine-for-indian-industry_amd"))
import mps  # noqa: E402

sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402

REF = "/root/reference/benchmarks"


def highs(lp):
    A = sp.csr_matrix((lp.val, lp.col, lp.row_ptr), shape=(lp.m, lp.n))
    le, ge, eq = lp.sense == 0, lp.sense == 1, lp.sense == 2
    A_ub = sp.vstack([A[le], -A[ge]]) if (le.any() or ge.any()) else None
    b_ub = np.concatenate([lp.rhs[le], -lp.rhs[ge]]) if A_ub is not None else None
    bounds = [(None if math.isinf(lo) else lo, None if math.isinf(hi) else hi) for lo, hi in zip(lp.lb, lp.ub)]
    c = -lp.obj if lp.maximize else lp.obj
    r = so.linprog(c, A_ub=A_ub, b_ub=b_ub, A_eq=A[eq] if eq.any() else None, b_eq=lp.rhs[eq] if eq.any() else None,
                   bounds=bounds, method="highs",
                   options={"primal_feasibility_tolerance": 1e-10, "dual_feasibility_tolerance": 1e-10})
    return r.status, (-r.fun if lp.maximize else r.fun) if r.status == 0 else None


def published():
    text = open(os.path.join(REF, "netlib_reference.cpp")).read()
    return {m.group(1): float(m.group(3)) for m in re.finditer(r'\{"(\w+)",\s*"(min|max)",\s*([-0-9.eE+]+)\}', text)}


def main():
    out = os.path.join(HERE, "netlib")
    os.makedirs(out, exist_ok=True)
    base = {r["name"]: r for r in csv.DictReader(open(os.path.join(REF, "results", "netlib-baseline-p0.csv")))}
    pub = published()
    table = {}
    for name in sorted(base):
        src = os.path.join(REF, "netlib", name + ".mps")
        dst = os.path.join(out, name + ".mps.gz")
        with open(src, "rb") as fi, gzip.GzipFile(dst, "wb", mtime=0) as fo:
            shutil.copyfileobj(fi, fo)
        lp, info = mps.read_mps(dst)
        st, opt = highs(lp)
        b = base[name]
        ref = O.ref_ipm_solve(lp, time_limit=30.0 if name == "dfl001" else 120.0)
        table[name] = {"ref_ipm_status": ref.status, "ref_ipm_iterations": int(ref.iterations),
                       "ref_ipm_objective": float(ref.objective),"cons": int(b["cons"]), "vars": int(b["vars"]), "nnz": int(b["nnz"]),
                       "published": pub.get(name), "highs": opt, "highs_status": int(st),
                       "reference_status": b["status"], "reference_objective": float(b["objective"]),
                       "reference_engine": b["engine"], "reference_time_s": float(b["time_s"])}
        print(name, table[name], flush=True)
    with open(os.path.join(HERE, "netlib.json"), "w") as f:
        json.dump(table, f, indent=1)


if __name__ == "__main__":
    main()
