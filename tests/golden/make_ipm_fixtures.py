"""Writes tests/golden/ipm_reference.json: what the reference's own
InteriorPointSolver (oracle/_ref, crossover off, industrial tolerances)
returns on the LPs of its IPM tests and on small seeded random LPs - status,
iteration count, objective, primal point and the per-iteration
HYPERNOVA_IPM_DBG trace (mu, rp, rd, ap, ad; printed with 6 significant
digits by the reference).

Run from the repository root where the reference build exists:
    python tests/golden/make_ipm_fixtures.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "indigenous-cpu-gpu-optimization-engine-for-indian-industry_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))
import lp_cases  # noqa: E402
import lpgen  # noqa: E402
import oracle as O  # noqa: E402


def record(lp, **kw):
    r = O.ref_ipm_solve(lp, **kw)
    return {"status": r.status, "iterations": r.iterations, "objective": r.objective,
            "primal": [float(v) for v in r.primal], "trace": r.trace.tolist(), "options": kw}


def main():
    out = {}
    for name, fn in lp_cases.REFERENCE_CASES.items():
        out[name] = record(fn())
    out["simple_lp_iter1"] = record(lp_cases.simple_lp(), max_iterations=1)   # test_lp.cpp:306-325
    out["simple_lp_interrupt"] = record(lp_cases.simple_lp(), interrupt=True)  # test_lp.cpp:287-304
    for seed in range(1, 6):
        out[f"random_60x150_s{seed}"] = record(lpgen.random_lp(seed, 60, 150))
    out["random_250x625_s7"] = record(lpgen.random_lp(7, 250, 625))
    out["transport_8x12_s0"] = record(lpgen.transport_lp(0, 8, 12))
    out["powerlaw_120x300_s3"] = record(lpgen.powerlaw_lp(3, 120, 300, 2.179, 2, 40))
    with open(os.path.join(HERE, "ipm_reference.json"), "w") as f:
        json.dump(out, f, indent=1)
    for k, v in out.items():
        print(k, v["status"], v["iterations"], v["objective"])


if __name__ == "__main__":
    main()
