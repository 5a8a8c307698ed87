"""Parity of the device verifier (hn_verify, through the C-ABI) with the
restated SolutionVerifier (oracle or_verify) and, where oracle/_ref is
present, with the reference's SolutionVerifier::verify_detailed itself:
every field bit-identical, the violated list and the message identical.
Covers the reference's own verifier tests (test_validation.cpp:14-161),
seeded problems over every branch, the IPM's results as api.cpp hands them
to the verifier, and the 200,000 x 500,000 power-law LP at full size."""
import numpy as np
import pytest

import lpgen
import oracle
import verify_cases as V
from lp_cases import simple_lp, larger_lp

pytestmark = pytest.mark.gpu
have_ref = oracle.have_reference() and hasattr(oracle.REF, "ref_verify")


def _gpu(hip, c: V.Case):
    return hip.verify(c.lp, c.primal, c.dual, c.reduced_costs, c.objective_value, c.var_type)


def _oracle(c: V.Case):
    return oracle.verify(c.lp, c.primal, c.dual, c.reduced_costs, c.objective_value, c.var_type, cap=1 << 16)


@pytest.mark.parametrize("case", V.unit_cases(), ids=lambda c: c.name)
def test_verify_reference_unit_tests(hip, case):
    print(f"  verify unit {case.name}", flush=True)
    r = _gpu(hip, case)
    V.check_unit_expectation(case, r)
    assert V.fields(r) == V.fields(_oracle(case))
    if have_ref:
        assert V.fields(r) == V.fields(oracle.ref_verify(case.lp, case.primal, case.dual, case.reduced_costs,
                                                         case.objective_value, case.var_type))


def test_verify_random_cases_match_oracle_and_reference(hip):
    cases = V.random_cases()
    for c in cases:
        print(f"  verify {c.name}", flush=True)
        r = V.fields(_gpu(hip, c))
        assert r == V.fields(_oracle(c)), c.name
        if have_ref:
            assert r == V.fields(oracle.ref_verify(c.lp, c.primal, c.dual, c.reduced_costs, c.objective_value,
                                                   c.var_type)), c.name


def test_verify_list_capacity_and_repeat_calls(hip):
    """A short list keeps the first entries and the full count; a smaller
    problem after a larger one reuses the device buffers correctly."""
    big = V.random_case(7, 300, 400, dual="full")
    small = V.random_case(8, 20, 30, dual="full")
    full = hip.SolutionVerifier().verify_detailed(big.lp, hip.Solution(big.primal, big.dual, None,
                                                                       big.objective_value))
    assert full.n_violated > 3
    short = hip.SolutionVerifier(max_listed=3).verify_detailed(big.lp, hip.Solution(big.primal, big.dual, None,
                                                                                    big.objective_value))
    assert short.n_violated == full.n_violated and short.violated_constraints == full.violated_constraints[:3]
    assert V.fields(_gpu(hip, small)) == V.fields(_oracle(small))
    assert V.fields(_gpu(hip, big)) == V.fields(_oracle(big))


def test_verify_size_mismatch_and_bad_input(hip):
    c = V.unit_cases()[0]
    r = hip.verify(c.lp, np.array([1.0]))
    assert not r.feasible and not r.optimal
    assert r.message == "Primal infeasible: Solution vector size mismatch"
    bad = V.from_rows(2, [[(0, 1.0), (5, 1.0)]], [V.LE], [1.0], [0, 0], [1, 1], [1.0, 1.0])
    with pytest.raises(hip.HnError):
        hip.verify(bad, np.array([0.5, 0.5]))


@pytest.mark.parametrize("make", [simple_lp, larger_lp], ids=["simple", "larger"])
def test_verify_gpu_ipm_results(hip, make):
    """The GPU IPM's result as api.cpp passes it on: primal, zero-filled dual
    (interior_point.cpp:771), objective."""
    lp = make()
    res = hip.ipm_solve(lp)
    for dual in (np.zeros(lp.m), None):
        c = V.Case("ipm", lp, res.primal, dual=dual, objective_value=res.objective_value)
        r = _gpu(hip, c)
        assert V.fields(r) == V.fields(_oracle(c))
        if dual is None:
            assert r.feasible and r.optimal, r.message


@pytest.mark.parametrize("which", ["a", "c"])
def test_verify_full_size(hip, which):
    """Config (a) 2,000 x 5,000 and config (c) 200,000 x 500,000 (power law,
    about 3M nonzeros): the generator's interior point, perturbed, with its
    dual, bit-identical to the restated verifier."""
    lp = lpgen.random_lp(1, 2000, 5000) if which == "a" else lpgen.powerlaw_lp(3, 200_000, 500_000)
    g = np.random.default_rng(5)
    x = lp.x0.copy()
    idx = g.choice(lp.n, size=lp.n // 100, replace=False)
    x[idx] = -g.uniform(1e-9, 1e-3, idx.size)
    objv = float(np.dot(lp.obj, x))
    for dual, rc in ((lp.y0, None), (lp.y0, lp.s0), (None, None)):
        c = V.Case(which, lp, x, dual=dual, reduced_costs=rc, objective_value=objv)
        print(f"  verify full {which} m={lp.m} n={lp.n} nnz={lp.nnz} dual={dual is not None} rc={rc is not None}",
              flush=True)
        r = _gpu(hip, c)
        o = _oracle(c)
        assert V.fields(r) == V.fields(o)
        assert r.n_violated >= idx.size
