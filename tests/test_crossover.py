"""Crossover from an IPM endpoint (the reference's simplex polish,
interior_point.cpp:786-878), CPU side: the reference's own crossover tests
(test_lp.cpp:327-400) on the reference built in oracle/_ref, the driver that
runs the polish from a given endpoint (ref_crossover) against the
reference's whole solve, and the restated IPM's endpoint polished to the
same vertex."""
import numpy as np
import pytest

import oracle
from lp_cases import crossover_cases, simple_lp

pytestmark = pytest.mark.skipif(not oracle.have_reference() or not hasattr(oracle.REF, "ref_crossover"),
                                reason="reference build not present")
CASES = crossover_cases()


@pytest.mark.parametrize("max_it,xo_it", [(1, 100), (100, 100), (1, 1)],
                         ids=["resolves_stall", "exact_certificate", "cold_fallback"])
def test_reference_crossover_tests(max_it, xo_it):
    """test_lp.cpp:327-400: OPTIMAL at 20, vertex (0, 10), duals (2, 0), b'y = obj."""
    r = oracle.ref_ipm_solve_crossover(simple_lp(), max_it, xo_it)
    assert r.status == "OPTIMAL"
    assert abs(r.objective - 20.0) <= 1e-6
    assert np.allclose(r.primal, [0.0, 10.0], atol=1e-6)
    assert np.allclose(r.dual, [2.0, 0.0], atol=1e-6)
    assert abs(10.0 * r.dual[0] + 5.0 * r.dual[1] - r.objective) <= 1e-6


@pytest.mark.parametrize("name,lp,max_it,xo_it", CASES, ids=[c[0] for c in CASES])
def test_polish_driver_matches_reference_solve(name, lp, max_it, xo_it):
    """The reference's barrier endpoint polished by ref_crossover is exactly
    the reference's own solve with crossover on."""
    full = oracle.ref_ipm_solve_crossover(lp, max_it, xo_it)
    bar = oracle.ref_ipm_solve(lp, max_iterations=max_it)
    xo = oracle.ref_crossover(lp, bar.status, bar.objective, bar.iterations, bar.primal, xo_it)
    assert (xo.status, xo.iterations) == (full.status, full.iterations)
    assert xo.objective == full.objective
    assert np.array_equal(xo.primal, full.primal) and np.array_equal(xo.dual, full.dual)


@pytest.mark.parametrize("name,lp,max_it,xo_it", CASES, ids=[c[0] for c in CASES])
def test_oracle_endpoint_polishes_to_reference_vertex(name, lp, max_it, xo_it):
    full = oracle.ref_ipm_solve_crossover(lp, max_it, xo_it)
    bar = oracle.ipm_solve(lp, max_iterations=max_it)
    xo = oracle.ref_crossover(lp, bar.status, bar.objective, bar.iterations, bar.primal, xo_it)
    assert xo.status == full.status
    assert abs(xo.objective - full.objective) <= 1e-9 * (1.0 + abs(full.objective))
    assert np.allclose(xo.primal, full.primal, rtol=1e-9, atol=1e-9)
    assert np.allclose(xo.dual, full.dual, rtol=1e-9, atol=1e-9)
