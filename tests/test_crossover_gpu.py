"""Crossover from the GPU IPM's endpoint: the GPU barrier (crossover off)
hands its primal to the reference's unchanged simplex polish
(interior_point.cpp:786-878, run by oracle/_ref's ref_crossover), which
must reach the vertex, duals and status of the reference's own solve with
crossover on; the polished (primal, dual) then goes through the device
verifier, bit-identical to the reference's SolutionVerifier."""
import numpy as np
import pytest

import oracle
from lp_cases import crossover_cases

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not oracle.have_reference() or not hasattr(oracle.REF, "ref_crossover"),
                                 reason="reference build not present")]
CASES = crossover_cases()


def _bits(r):
    return [np.float64(v).view(np.uint64).item() for v in
            (r.primal_infeasibility, r.dual_infeasibility, r.complementarity, r.integrality_violation,
             r.objective_discrepancy)] + [r.feasible, r.optimal, r.n_violated, r.message]


@pytest.mark.parametrize("name,lp,max_it,xo_it", CASES, ids=[c[0] for c in CASES])
def test_gpu_endpoint_crossover_matches_reference_solve(hip, name, lp, max_it, xo_it):
    print(f"  crossover {name} m={lp.m} n={lp.n} max_it={max_it} xo_it={xo_it}", flush=True)
    full = oracle.ref_ipm_solve_crossover(lp, max_it, xo_it)
    bar = hip.ipm_solve(lp, max_iterations=max_it)
    xo = oracle.ref_crossover(lp, bar.status, bar.objective_value, bar.iterations, bar.primal, xo_it)
    assert xo.status == full.status
    assert abs(xo.objective - full.objective) <= 1e-9 * (1.0 + abs(full.objective))
    assert np.allclose(xo.primal, full.primal, rtol=1e-9, atol=1e-9)
    assert np.allclose(xo.dual, full.dual, rtol=1e-9, atol=1e-9)
    # the api's post-solve check on the polished point (api.cpp:510,523)
    g = hip.verify(lp, xo.primal, xo.dual, None, xo.objective)
    ref = oracle.ref_verify(lp, xo.primal, xo.dual, None, xo.objective)
    assert _bits(g) == _bits(ref)
    assert g.violated_constraints == ref.violated
    assert g.feasible and g.optimal, g.message
