"""Inputs for the post-solve verifier tests: the reference's own
SolutionVerifier unit tests (tests/unit/test_validation.cpp:14-161) and
seeded random problems whose solutions hit every branch of
core/validation/solution_verifier.cpp (bound and row violations of each
sense, rows within and just beyond the tolerance, NaN and Inf values,
integer columns, given or computed reduced costs, duals of the wrong
length, maximisation)."""
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

import lpgen
from lp_cases import from_rows

INF = np.inf
LE, GE, EQ = lpgen.LE, lpgen.GE, lpgen.EQ
CONT, INT, BIN = 0, 1, 2


@dataclass
class Case:
    name: str
    lp: lpgen.LP
    primal: np.ndarray
    dual: Optional[np.ndarray] = None
    reduced_costs: Optional[np.ndarray] = None
    objective_value: float = 0.0
    var_type: Optional[np.ndarray] = None
    expect: dict = field(default_factory=dict)  # what the reference's test asserts


def unit_cases():
    """test_validation.cpp:14-161, one Case per TEST."""
    one_le5 = lambda: from_rows(1, [[(0, 1.0)]], [LE], [5.0], [0], [10], [1.0], maximize=True)
    bare = lambda ub, c=1.0: from_rows(1, [], [], [], [0], [ub], [c], maximize=True)
    return [
        Case("FeasibleSolution", from_rows(2, [[(0, 1.0), (1, 1.0)]], [LE], [10.0], [0, 0], [10, 10], [1.0, 2.0],
                                           maximize=True),
             np.array([5.0, 5.0]), dual=np.array([1.0]), expect=dict(feasible=True)),
        Case("InfeasiblePrimal", one_le5(), np.array([10.0]), expect=dict(verify=False)),
        Case("InfeasibleDual", one_le5(), np.array([5.0]), dual=np.array([100.0]), expect=dict(optimal=False)),
        Case("IntegralityCheck", bare(10.0), np.array([5.5]), var_type=np.array([INT], np.int8),
             expect=dict(feasible=False, integrality_positive=True)),
        Case("IntegerFeasible", bare(10.0), np.array([5.0]), var_type=np.array([INT], np.int8),
             expect=dict(feasible=True)),
        Case("BoundViolation", bare(5.0), np.array([10.0]), expect=dict(feasible=False, primal_positive=True)),
        Case("NaNInSolution", bare(10.0), np.array([np.nan]), expect=dict(feasible=False)),
        Case("ObjectiveDiscrepancy", bare(10.0), np.array([5.0]), objective_value=100.0,
             expect=dict(optimal=False, discrepancy_positive=True)),
        Case("ObjectiveMatches", bare(10.0, 2.0), np.array([5.0]), objective_value=10.0,
             expect=dict(discrepancy_zero=True)),
    ]


def check_unit_expectation(case: Case, r) -> None:
    """The reference test's assertions on a VerificationResult-like r."""
    e = case.expect
    if "feasible" in e:
        assert r.feasible == e["feasible"], case.name
    if "optimal" in e:
        assert r.optimal == e["optimal"], case.name
    if "verify" in e:
        assert (r.feasible and r.optimal) == e["verify"], case.name
    if e.get("integrality_positive"):
        assert r.integrality_violation > 0.0, case.name
    if e.get("primal_positive"):
        assert r.primal_infeasibility > 0.0, case.name
    if e.get("discrepancy_positive"):
        assert r.objective_discrepancy > 0.0, case.name
    if e.get("discrepancy_zero"):
        assert r.objective_discrepancy == 0.0, case.name


def _activity(lp, x):
    """Row activities summed left to right (the verifier's order)."""
    out = np.zeros(lp.m)
    for i in range(lp.m):
        s = 0.0
        for k in range(lp.row_ptr[i], lp.row_ptr[i + 1]):
            s += float(lp.val[k]) * float(x[lp.col[k]])
        out[i] = s
    return out


def random_case(seed: int, m: int, n: int, per_row: int = 6, nonfinite: bool = False, integer: bool = False,
                dual: str = "full", rc: bool = False, maximize: Optional[bool] = None, clean: bool = False) -> Case:
    """A seeded problem and solution.  dual: "full" (length m), "none",
    "short" (length m - 1, treated as absent), "zero" (the IPM's zero-filled
    dual).  clean: a primal within its bounds and rows, reduced costs of
    1e-12 and the exact objective, so only complementarity can fail."""
    g = np.random.default_rng(seed)
    rows = []
    for i in range(m):
        k = int(g.integers(0, per_row * 2 + 1)) if i % 7 else 0  # some empty rows
        cols = np.sort(g.choice(n, size=min(k, n), replace=False))
        scale = 10.0 ** g.uniform(-2, 2)
        rows.append([(int(j), float(g.normal() * scale)) for j in cols])
    lb = np.where(g.random(n) < 0.7, 0.0, np.where(g.random(n) < 0.5, -INF, g.normal(size=n)))
    ub = np.where(g.random(n) < 0.6, INF, lb + np.abs(g.normal(size=n)) * 5)
    ub = np.where(np.isinf(lb) & np.isinf(ub), INF, ub)
    fixed = g.random(n) < 0.05
    ub = np.where(fixed & np.isfinite(lb), lb, ub)
    x = np.where(np.isfinite(lb), lb, 0.0) + np.where(g.random(n) < 0.5, 0.0, g.uniform(0, 3, n))
    x = np.minimum(x, np.where(np.isfinite(ub), ub, x))
    # bound violations, some below the tolerance and some beyond
    for j in ([] if clean else g.choice(n, size=max(1, n // 20), replace=False)):
        tiny = g.random() < 0.5
        d = (1e-10 if tiny else 10.0 ** g.uniform(-6, 1)) * (1 + abs(x[j]))
        if g.random() < 0.5 and np.isfinite(lb[j]):
            x[j] = lb[j] - d
        elif np.isfinite(ub[j]):
            x[j] = ub[j] + d
    obj = g.normal(size=n) * np.where(g.random(n) < 0.1, 0.0, 1.0)
    mx = bool(g.random() < 0.5) if maximize is None else maximize
    senses = g.integers(0, 3, size=m)
    lp = from_rows(n, rows, senses.tolist(), np.zeros(m).tolist(), lb, ub, obj, maximize=mx)
    lp.obj_offset = float(g.normal()) if g.random() < 0.5 else 0.0
    act = _activity(lp, x)
    absact = np.array([sum(abs(float(lp.val[k]) * float(x[lp.col[k]]))
                           for k in range(lp.row_ptr[i], lp.row_ptr[i + 1])) for i in range(m)])
    rhs = act.copy()
    ftol = 1e-9 * 10.0
    for i in range(m):
        u = g.random() * (0.35 if clean else 1.0)
        scale = 1.0 + abs(act[i]) + absact[i]
        if u < 0.35:    # satisfied with slack (for LE/GE), exact for EQ
            slack = abs(g.normal()) * scale
            rhs[i] = act[i] + (slack if senses[i] == LE else (-slack if senses[i] == GE else 0.0))
        elif u < 0.55:  # violated clearly
            d = 10.0 ** g.uniform(-5, 1) * scale
            rhs[i] = act[i] + (-d if senses[i] == LE else d)
        elif u < 0.8:   # near the tolerance, either side
            rhs[i] = act[i] - (1 if senses[i] != GE else -1) * ftol * scale * (1 + g.choice([-1e-3, 1e-3, 0.0]))
        # else: exactly the activity
    lp.rhs = rhs
    if nonfinite:
        j = int(g.integers(n // 3, n))
        x[j] = g.choice([np.nan, np.inf, -np.inf])
    vt = None
    if integer:
        vt = g.choice([CONT, INT, BIN, 3], size=n, p=[0.6, 0.2, 0.15, 0.05]).astype(np.int8)
        for j in np.nonzero(vt > 0)[0][::2]:
            if np.isfinite(x[j]):
                x[j] = np.round(x[j]) + (g.choice([0.0, 1e-7, 1e-5, 0.3]))
    y = None
    if dual == "full":
        y = g.normal(size=m) * np.where(g.random(m) < 0.3, 0.0, 1.0) * 10.0 ** g.uniform(-3, 2, m)
    elif dual == "short":
        y = g.normal(size=max(m - 1, 0))
    elif dual == "zero":
        y = np.zeros(m)
    r = None
    if rc:
        r = g.normal(size=n) * np.where(g.random(n) < 0.5, 1e-12, 1.0)
    if clean:
        r = np.full(n, 1e-12)
    objv = lp.obj_offset
    for j in range(n):
        objv += float(obj[j]) * float(x[j])
    if g.random() < 0.3 and not clean:
        objv += g.choice([1e-12, 1.0, 100.0])
    return Case(f"rand{seed}_{m}x{n}", lp, x, dual=y, reduced_costs=r, objective_value=objv, var_type=vt)


def random_cases():
    out = []
    s = 100
    for m, n in ((1, 1), (3, 5), (20, 30), (50, 40), (200, 500), (300, 120)):
        for nonfinite in (False, True):
            for integer in (False, True):
                for dual in ("full", "none", "short", "zero"):
                    rc = dual == "full" and (s % 3 == 0)
                    out.append(random_case(s, m, n, nonfinite=nonfinite, integer=integer, dual=dual, rc=rc))
                    s += 1
    for m, n in ((5, 8), (40, 60), (200, 300)):
        for dual in ("full", "none"):
            out.append(random_case(s, m, n, dual=dual, clean=True))
            s += 1
    return out


def fields(r):
    """The comparable fields of a verifier result, floats by their bits."""
    def bits(v):
        return np.float64(v).view(np.uint64).item()
    return dict(feasible=bool(r.feasible), optimal=bool(r.optimal),
                primal=bits(r.primal_infeasibility), dual=bits(r.dual_infeasibility),
                comp=bits(r.complementarity), integ=bits(r.integrality_violation),
                disc=bits(r.objective_discrepancy), n_violated=int(r.n_violated),
                violated=list(getattr(r, "violated", None) or getattr(r, "violated_constraints", [])),
                message=r.message)
