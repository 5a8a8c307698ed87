"""KKT measures on the caller's own LP (hn_ipm_result.rel_*), recomputed here
in numpy from the returned point.

The library maps the scaled standard-form point back (x by to_original_primal,
interior_point.cpp:375-395; row duals gamma_i lam_r; bound multipliers
s_k / k_k) and measures ||Ax - b||, ||A'y + z - c|| and |c'x - b'y| on the
caller's data.  This file restates the map for the standard-form
workloads, where it is a diagonal rescaling fixed by the reference's row
scale 1/max_j |a_ij| and column scale 1/max_i |m_ij|
(interior_point.cpp:282-291,345-370), and for the transportation LP (LE
supply rows with slacks).  A general LP with every bound type is pinned by
its optimum: at a converged point the mapped duals must be dual feasible
and close the gap with HiGHS's optimal value."""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.optimize import linprog

import lp_cases
import lpgen

pytestmark = pytest.mark.gpu


def _csr(lp):
    return sp.csr_matrix((lp.val, lp.col, lp.row_ptr), shape=(lp.m, lp.n))


def kkt_standard(lp, r):
    """All rows EQ, 0 <= x: y = G lam, z = s / K."""
    A = _csr(lp)
    G = 1.0 / abs(A).max(axis=1).toarray().ravel()
    K = 1.0 / abs(sp.diags(G) @ A).max(axis=0).toarray().ravel()
    x = r.primal
    y = G * r.y
    z = r.s / K
    c, b = lp.obj, lp.rhs
    rp = np.max(np.abs(A @ x - b)) / (1 + np.max(np.abs(b)))
    rd = np.max(np.abs(A.T @ y + z - c)) / (1 + np.max(np.abs(c)))
    cx = c @ x
    gap = abs(cx - b @ y) / (1 + abs(cx))
    return rp, rd, gap


def kkt_transport(lp, r):
    """S LE supply rows (one slack each, in row order) then EQ demand rows;
    every coefficient is 1, so the row and column scales are 1."""
    A = _csr(lp)
    assert np.all(lp.val == 1.0)
    le = np.flatnonzero(lp.sense == lpgen.LE)
    x, y = r.primal, r.y[:lp.m]
    z = r.s[:lp.n]
    w = r.s[lp.n:lp.n + le.size]
    c, b = lp.obj, lp.rhs
    viol = A @ x - b
    viol[le] = np.maximum(0.0, viol[le])
    rp = np.max(np.abs(viol)) / (1 + np.max(np.abs(b)))
    rd = max(np.max(np.abs(A.T @ y + z - c)), np.max(np.abs(y[le] + w))) / (1 + np.max(np.abs(c)))
    cx = c @ x
    gap = abs(cx - b @ y) / (1 + abs(cx))
    return rp, rd, gap


def close(lib, ref):
    return abs(lib - ref) <= 1e-12 + 1e-6 * abs(ref)


def _check(r, ref, what):
    got = (r.rel_primal, r.rel_dual, r.rel_gap)
    print(f"[kkt] {what}: library {got}, numpy {ref}", flush=True)
    for g, w in zip(got, ref):
        assert close(g, w), (what, got, ref)


def test_single_lp_metrics_match_numpy(hip):
    lp = lpgen.random_lp(1, 2000, 5000)
    print("[case] kkt single random 2000x5000 seed 1, gap_tol 1e-8", flush=True)
    r = hip.ipm_solve(lp, gap_tol=1e-8)
    assert r.status == "OPTIMAL"
    ref = kkt_standard(lp, r)
    _check(r, ref, "2000x5000")
    assert max(ref) <= 1e-8
    # the standard form here has no bound shifts: its gap is the caller's
    assert abs(r.scaled_rel_gap - r.rel_gap) <= 1e-12


def test_batch_metrics_match_numpy(hip):
    lps = [lpgen.batch_lp(i) for i in range(0, 4096, 256)]
    print(f"[case] kkt batch of {len(lps)} (m {min(l.m for l in lps)}..{max(l.m for l in lps)}), skip rule",
          flush=True)
    res = hip.ipm_solve_batch(lps, want_points=True, pivot_rule=hip.HN_PIVOT_SKIP, gap_tol=1e-8)
    for i, (lp, r) in enumerate(zip(lps, res)):
        assert r.status == "OPTIMAL"
        ref = kkt_standard(lp, r)
        _check(r, ref, f"batch {i} m={lp.m}")
        assert max(ref) <= 1e-8


def test_transport_metrics_match_numpy(hip):
    lp = lpgen.transport_lp(1_000_000)
    print("[case] kkt transport 200x300 seed 1000000, skip rule", flush=True)
    r = hip.ipm_solve(lp, pivot_rule=hip.HN_PIVOT_SKIP)
    assert r.status == "OPTIMAL"
    _check(r, kkt_transport(lp, r), "transport")


GENERAL = {"general_s0": lambda: lp_cases.general_bounds_lp(0), "general_s1": lambda: lp_cases.general_bounds_lp(1),
           "general_max_s2": lambda: lp_cases.general_bounds_lp(2, maximize=True),
           "simple_lp": lp_cases.simple_lp, "integration_lp": lp_cases.integration_lp,
           "gpu_parity_lp": lp_cases.gpu_parity_lp}


@pytest.mark.parametrize("name", sorted(GENERAL))
def test_general_lp_metrics_close_the_gap(hip, name):
    """Fixed, lb-only, ub-only and boxed columns, LE/GE/EQ rows, a duplicate
    row, an objective offset, maximisation: at a converged point the mapped
    duals are dual feasible and the gap to HiGHS's optimum closes.  A wrong
    sign or scale in the map (e.g. a GE row not negated, a bound multiplier
    not divided by its column scale) leaves an O(1) dual residual or gap."""
    lp = GENERAL[name]()
    print(f"[case] kkt general {name}", flush=True)
    r = hip.ipm_solve(lp, pivot_rule=hip.HN_PIVOT_SKIP, gap_tol=1e-9)
    A = _csr(lp)
    le, ge, eq = (lp.sense == lpgen.LE), (lp.sense == lpgen.GE), (lp.sense == lpgen.EQ)
    c = -lp.obj if lp.maximize else lp.obj
    h = linprog(c, A_ub=sp.vstack([A[le], -A[ge]]) if (le.any() or ge.any()) else None,
                b_ub=np.concatenate([lp.rhs[le], -lp.rhs[ge]]) if (le.any() or ge.any()) else None,
                A_eq=A[eq] if eq.any() else None, b_eq=lp.rhs[eq] if eq.any() else None,
                bounds=list(zip(lp.lb, lp.ub)), method="highs")
    assert h.status == 0
    print(f"[kkt] {name}: status {r.status} rel {r.rel_primal:.2e} {r.rel_dual:.2e} {r.rel_gap:.2e}, "
          f"obj {r.objective_value!r} vs HiGHS {(-h.fun if lp.maximize else h.fun)!r}", flush=True)
    assert r.status == "OPTIMAL"
    assert r.rel_primal <= 1e-8 and r.rel_dual <= 1e-8 and r.rel_gap <= 1e-9
    best = (-h.fun if lp.maximize else h.fun) + lp.obj_offset
    assert abs(best - r.objective_value) <= 1e-7 * (1 + abs(best))
