"""The restated post-solve verifier (oracle or_verify) pinned to the
reference: the assertions of the reference's own SolutionVerifier tests
(tests/unit/test_validation.cpp:14-161) and, where oracle/_ref is built, the
reference's SolutionVerifier::verify_detailed itself, field by field and bit
by bit, on those tests, on seeded problems covering every branch and on the
reference IPM's own results."""
import numpy as np
import pytest

import oracle
import verify_cases as V
from lp_cases import simple_lp, larger_lp

need_ref = pytest.mark.skipif(not oracle.have_reference() or not hasattr(oracle.REF, "ref_verify"),
                              reason="reference build not present")


def _oracle(c: V.Case):
    return oracle.verify(c.lp, c.primal, c.dual, c.reduced_costs, c.objective_value, c.var_type)


def _ref(c: V.Case):
    return oracle.ref_verify(c.lp, c.primal, c.dual, c.reduced_costs, c.objective_value, c.var_type)


@pytest.mark.parametrize("case", V.unit_cases(), ids=lambda c: c.name)
def test_oracle_meets_reference_unit_tests(case):
    V.check_unit_expectation(case, _oracle(case))


@need_ref
@pytest.mark.parametrize("case", V.unit_cases(), ids=lambda c: c.name)
def test_oracle_matches_reference_on_unit_tests(case):
    assert V.fields(_oracle(case)) == V.fields(_ref(case))


@need_ref
@pytest.mark.parametrize("case", V.random_cases(), ids=lambda c: c.name)
def test_oracle_matches_reference_random(case):
    a, b = V.fields(_oracle(case)), V.fields(_ref(case))
    assert a == b


def test_random_cases_cover_the_branches():
    """The random set is not vacuous: it holds passing and failing results of
    every kind."""
    rs = [_oracle(c) for c in V.random_cases()]
    msgs = {r.message.split(":")[0] for r in rs}
    assert {"Primal infeasible", "Dual infeasible", "Complementarity violation", "Integrality violation",
            "Objective discrepancy"} <= msgs
    assert any(r.feasible and r.optimal for r in rs)
    assert any("NaN or Inf" in r.message for r in rs)
    assert any(r.n_violated > 1 for r in rs)


@need_ref
@pytest.mark.parametrize("make", [simple_lp, larger_lp], ids=["simple", "larger"])
def test_oracle_matches_reference_on_ipm_results(make):
    """The reference IPM's own result as api.cpp hands it to the verifier:
    its primal, a zero-filled dual (interior_point.cpp:771), its objective."""
    lp = make()
    r = oracle.ref_ipm_solve(lp)
    for dual in (np.zeros(lp.m), None):
        c = V.Case("ipm", lp, r.primal, dual=dual, objective_value=r.objective)
        assert V.fields(_oracle(c)) == V.fields(_ref(c))


def test_size_mismatch():
    c = V.unit_cases()[0]
    r = oracle.verify(c.lp, np.array([1.0]), None)
    assert not r.feasible and not r.optimal
    assert r.message == "Primal infeasible: Solution vector size mismatch"


def test_verifier_messages_follow_priority():
    """Integrality overrides the primal message; the dual message is only set
    while feasible (solution_verifier.cpp:32-67)."""
    lp = V.from_rows(1, [[(0, 1.0)]], [V.LE], [1.0], [0], [10], [1.0])
    r = oracle.verify(lp, np.array([2.5]), np.array([5.0]), var_type=np.array([V.INT], np.int8),
                      objective_value=2.5)
    assert r.message.startswith("Integrality violation: Max integrality violation: 0.500000")
    r = oracle.verify(lp, np.array([0.5]), np.array([5.0]), objective_value=0.5)
    assert r.message.startswith("Complementarity violation")
