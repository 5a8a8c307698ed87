"""The GPU IPM on the reference's 21 Netlib models (tests/golden/netlib/),
against the outcome of the reference's own InteriorPointSolver on the same
parsed LPs (tests/golden/netlib.json, recorded from oracle/_ref by
tests/golden/make_netlib_fixtures.py) and against HiGHS.

Under the reference's pivot rule the GPU solver must reproduce the
reference model by model: same status and iteration count; the objective
within 1e-8 where the reference converges (fewer than 100 iterations) and
within 1e-3 where it stalls at the iteration limit (the end game of a
stalled solve is not reproducible to rounding).  Two models differ by
design: bandm, where the reference's minimum-degree sparse LDL' ends in
NUMERICAL_ERROR at iteration 18 and the dense factor here carries on, and
dfl001, where the reference does not finish its ordering in 30 s.
"""
import json
import os

import pytest

import mps

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = json.load(open(os.path.join(HERE, "golden", "netlib.json")))


def _lp(name):
    lp, _ = mps.read_mps(os.path.join(HERE, "golden", "netlib", name + ".mps.gz"))
    return lp


def rel(a, b):
    return abs(a - b) / max(1.0, abs(b))


@pytest.mark.parametrize("name", sorted(n for n in FIX if n not in ("bandm", "dfl001")))
def test_netlib_matches_reference_ipm(hip, name):
    d = FIX[name]
    print(f"[case] netlib {name} (reference {d['ref_ipm_status']} in {d['ref_ipm_iterations']})", flush=True)
    r = hip.ipm_solve(_lp(name))
    assert r.status == d["ref_ipm_status"], (r.status, d["ref_ipm_status"])
    assert r.iterations == d["ref_ipm_iterations"]
    tol = 1e-8 if d["ref_ipm_iterations"] < 100 else 1e-3
    assert rel(r.objective_value, d["ref_ipm_objective"]) <= tol
    if d["ref_ipm_iterations"] < 100:  # converged: the optimum itself
        assert rel(r.objective_value, d["highs"]) <= 1e-7


def test_netlib_bandm_dense_factor_carries_on(hip):
    d = FIX["bandm"]
    assert d["ref_ipm_status"] == "NUMERICAL_ERROR" and d["ref_ipm_iterations"] == 18
    r = hip.ipm_solve(_lp("bandm"))
    assert r.status == "OPTIMAL" and rel(r.objective_value, d["highs"]) <= 1e-6


def test_netlib_dfl001(hip):
    # 6,071 x 12,230: the reference finishes no iteration in 30 s; the GPU
    # solver runs its 100 iterations and ends near the HiGHS optimum
    d = FIX["dfl001"]
    assert d["ref_ipm_status"] == "TIME_LIMIT" and d["ref_ipm_iterations"] == 0
    print("[case] netlib dfl001", flush=True)
    r = hip.ipm_solve(_lp("dfl001"))
    assert r.status in ("OPTIMAL", "ITER_LIMIT") and r.iterations <= 100
    assert rel(r.objective_value, d["highs"]) <= 1e-5
