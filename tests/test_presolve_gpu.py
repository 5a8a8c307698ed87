"""Presolve in front of the GPU IPM, as Solver::solve runs it
(api/cpp/api.cpp:384-412, then :176-196): the reference's Presolver
(oracle/_ref, AGGRESSIVE, the default level) reduces the problem, the GPU
IPM solves the reduced problem under the reference's rules, the reference's
postsolve maps the point back and the objective is recomputed on the
original.  Checked against the restated IPM on the same reduced problem
(status, iterations, objective), HiGHS's optimum of the original (Netlib
fixtures) and the device verifier's primal checks."""
import json
import os

import numpy as np
import pytest

import lpgen
import mps
import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not oracle.have_reference() or not hasattr(oracle.REF, "ref_presolve"),
                                 reason="reference build not present")]
FX = json.load(open(os.path.join(GOLDEN, "netlib.json")))
# models whose reduced form the restated (dense, natural order) IPM solves in seconds
SMALL = [n for n in sorted(FX) if n not in ("25fv47", "dfl001", "sctap2", "sctap3")]


@pytest.mark.parametrize("name", SMALL)
def test_presolved_netlib_on_gpu(hip, name):
    lp = mps.read_mps(os.path.join(GOLDEN, "netlib", name + ".mps.gz"))[0]
    P = oracle.RefPresolve(lp)
    red = P.reduced
    print(f"  presolve {name}: {lp.m}x{lp.n} -> {red.m}x{red.n} in {P.ms:.2f} ms", flush=True)
    g = hip.ipm_solve(red)
    o = oracle.ipm_solve(red)
    x, _, obj, _ = P.postsolve(g.primal)
    highs = FX[name]["highs"]
    if o.iterations < 100:  # the restated reference converges: the GPU must match it
        assert (g.status, g.iterations) == (o.status, o.iterations)
        assert abs(g.objective_value - o.objective) <= 1e-9 * (1.0 + abs(o.objective))
        assert abs(obj - highs) <= 1e-6 * max(1.0, abs(highs))
        v = hip.verify(lp, x, None, None, obj)
        assert v.feasible, v.message
    else:  # the restated reference stalls; rounding decides whether the GPU stalls too
        assert g.iterations == 100 or abs(obj - highs) <= 1e-6 * max(1.0, abs(highs))


@pytest.mark.parametrize("which", ["a", "transport"])
def test_presolve_pipeline_full_size(hip, which):
    """2,000 x 5,000 LP and the 200 x 300 transportation LP: nothing to remove,
    so presolve -> GPU IPM -> postsolve gives the direct GPU solve's bits."""
    lp = lpgen.random_lp(1, 2000, 5000) if which == "a" else lpgen.transport_lp(1_000_000, 200, 300)
    P = oracle.RefPresolve(lp)
    red = P.reduced
    print(f"  presolve {which}: {lp.m}x{lp.n} -> {red.m}x{red.n} in {P.ms:.2f} ms", flush=True)
    assert (red.m, red.n, red.nnz) == (lp.m, lp.n, lp.nnz)
    direct = hip.ipm_solve(lp)
    g = hip.ipm_solve(red)
    x, _, obj, _ = P.postsolve(g.primal)
    assert (g.status, g.iterations) == (direct.status, direct.iterations)
    assert np.array_equal(x, direct.primal)
    assert obj == direct.objective_value
