"""Sparse LDL' on the device (hn_sldl_factor / hn_sldl_solve_dev) and the
IPM's sparse normal-equation solver (HN_NE_SPARSE).

Bars:
  * L, D and the solve bit-identical to the reference's own SparseCholesky
    (oracle/_ref) on the same CSR matrix: normal equations M D M' + r I of
    real LPs with random positive D, the reference's unit-test matrices, and
    a rank-deficient N whose zero pivot is clamped to 1e-10 (the clamp count
    matches the reference's singular flag); on small cases also identical to
    the oracle's restatement;
  * the IPM with the sparse solver reproduces the reference IPM model by
    model on the reference's Netlib set (status and iteration count, as
    tests/test_netlib_gpu.py does for the dense solver)."""
import json
import os

import numpy as np
import pytest

import lpgen
import mps
import oracle as O

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
need_ref = pytest.mark.skipif(not O.have_reference() or not hasattr(O.REF, "ref_chol_create"),
                              reason="oracle/_ref (the reference build) not present")


def _netlib(name):
    lp, _ = mps.read_mps(os.path.join(HERE, "golden", "netlib", name + ".mps.gz"))
    return lp


def _normal_eq(lp, seed, reg=1e-8):
    """N = M D M' + reg I as a CSR with the normal-equation pattern (rows
    ascending), D > 0 random."""
    sf = O.standard_form(lp)
    m, n = sf["m"], sf["n"]
    rng = np.random.default_rng(seed)
    D = np.exp(rng.uniform(-6, 6, n))
    rp, ci = O.adat_pattern(sf["rp"], sf["ci"], m, n)
    Nd, _ = O.adat_dense(sf["rp"], sf["ci"], sf["v"], D, m, n, reg)
    v = np.array([Nd[i, ci[k]] for i in range(m) for k in range(rp[i], rp[i + 1])])
    return m, rp, ci, v


def _spd_from_pattern(n, rp, ci, seed):
    rng = np.random.default_rng(seed)
    A = np.zeros((n, n))
    for i in range(n):
        for k in range(rp[i], rp[i + 1]):
            j = ci[k]
            if j < i:
                A[i, j] = A[j, i] = rng.uniform(-1, 1)
    A += np.diag(np.abs(A).sum(1) + 1.0)
    return np.array([A[i, ci[k]] for i in range(n) for k in range(rp[i], rp[i + 1])])


CASES = {
    "cycle4": lambda: (4, np.array([0, 3, 6, 9, 12]), np.array([0, 1, 2, 0, 1, 3, 0, 2, 3, 1, 2, 3]),
                       np.array([4.0, 1, 1, 1, 4, 1, 1, 4, 1, 1, 1, 4])),
    "afiro": lambda: _normal_eq(_netlib("afiro"), 1),
    "blend": lambda: _normal_eq(_netlib("blend"), 2),
    "sctap1": lambda: _normal_eq(_netlib("sctap1"), 3),
    "bandm": lambda: _normal_eq(_netlib("bandm"), 4),
    "share1b": lambda: _normal_eq(_netlib("share1b"), 5),
    "transport_30x40": lambda: _normal_eq(lpgen.transport_lp(2, 30, 40), 6),
    "random_lp_200x500": lambda: _normal_eq(lpgen.random_lp(3, 200, 500), 7),
}


def _check_factor(hip, n, rp, ci, v, small):
    A = hip.CsrMatrix.upload(v, ci, rp, n, hip.HN_F64)
    s = hip.SparseLDL(rp, ci)
    clamped = s.factor(A)
    L, D = s.factors()
    R = O.ref_sparse_cholesky(rp, ci, v, b=np.linspace(-1.0, 2.0, n))
    assert np.array_equal(L, R["L"]), np.abs(L - R["L"]).max()
    assert np.array_equal(D, R["D"])
    assert (clamped > 0) == R["singular"]
    x = s.solve(np.linspace(-1.0, 2.0, n))
    assert np.array_equal(x, R["x"]), np.abs(x - R["x"]).max()
    if small:
        sy = s.symbolic()
        Lo, Do, sing = O.sparse_ldl(rp, ci, v, sy["perm"], sy["l_row_ptr"], sy["l_col"])
        assert np.array_equal(L, Lo) and np.array_equal(D, Do) and sing == R["singular"]
        xo = O.sparse_ldl_solve(sy["perm"], sy["l_row_ptr"], sy["l_col"], Lo, Do, np.linspace(-1.0, 2.0, n))
        assert np.array_equal(x, xo)
    return clamped


@need_ref
@pytest.mark.parametrize("name", sorted(CASES))
def test_sparse_ldl_bit_identical_to_reference(hip, name):
    n, rp, ci, v = CASES[name]()
    print(f"[case] sparse LDL' {name} n={n} nnz={len(ci)}", flush=True)
    _check_factor(hip, n, rp, ci, v, small=n <= 80)


@need_ref
def test_sparse_ldl_random_spd_and_refactorize(hip):
    # random SPD values on real patterns, and a second factorisation of the
    # same handle with new values (SparseCholeskyRefactorizeReuse)
    m, rp, ci, _ = _normal_eq(_netlib("sctap1"), 8)
    A = hip.CsrMatrix.upload(_spd_from_pattern(m, rp, ci, 1), ci, rp, m, hip.HN_F64)
    s = hip.SparseLDL(rp, ci)
    for seed in (1, 2, 3):
        v = _spd_from_pattern(m, rp, ci, seed)
        A = hip.CsrMatrix.upload(v, ci, rp, m, hip.HN_F64)
        assert s.factor(A) == 0
        L, D = s.factors()
        R = O.ref_sparse_cholesky(rp, ci, v, b=np.ones(m))
        assert np.array_equal(L, R["L"]) and np.array_equal(D, R["D"])
        assert np.array_equal(s.solve(np.ones(m)), R["x"])


@need_ref
def test_sparse_ldl_clamps_a_zero_pivot(hip):
    # two identical rows of M make N singular; with only the reference's
    # relative regularisation (1e-12 max diag N) one pivot falls below 1e-10
    # and is replaced by 1e-10, as the reference does
    lp = lpgen.random_lp(5, 30, 80)
    sf = O.standard_form(lp)
    m, n = sf["m"], sf["n"]
    rp, ci, v = sf["rp"], sf["ci"], sf["v"]
    rp2 = np.concatenate([rp, [rp[-1] + rp[1] - rp[0]]])
    ci2 = np.concatenate([ci, ci[rp[0]:rp[1]]])
    v2 = np.concatenate([v, v[rp[0]:rp[1]]])
    D = np.ones(n)
    prp, pci = O.adat_pattern(rp2, ci2, m + 1, n)
    Nd, _ = O.adat_dense(rp2, ci2, v2, D, m + 1, n, 0.0)
    pv = np.array([Nd[i, pci[k]] for i in range(m + 1) for k in range(prp[i], prp[i + 1])])
    clamped = _check_factor(hip, m + 1, prp, pci, pv, small=True)
    assert clamped >= 1


FIX = json.load(open(os.path.join(HERE, "golden", "netlib.json")))


# The models whose factor has short rows (each solve well under a second);
# 25fv47, israel, sctap2, sctap3, shell, stair and tuff also match status and
# iteration count but take 2-57 s each here (long dense rows of L are a
# sequential chain per row).
SPARSE_FAST = ["adlittle", "afiro", "blend", "kb2", "recipe", "sc105", "sc205", "sc50a", "sc50b", "sctap1",
               "share1b", "share2b"]


@pytest.mark.parametrize("name", SPARSE_FAST)
def test_ipm_sparse_solver_matches_reference_ipm(hip, name):
    d = FIX[name]
    print(f"[case] netlib {name} sparse solver (reference {d['ref_ipm_status']} in {d['ref_ipm_iterations']})",
          flush=True)
    r = hip.ipm_solve(_netlib(name), solver=hip.HN_NE_SPARSE)
    print(f"  {r.status} in {r.iterations}, obj {r.objective_value:.10e}, nnz(L) {r.sparse_l_nnz}, "
          f"{r.solve_time_ms:.1f} ms", flush=True)
    assert r.sparse_l_nnz > 0
    assert r.status == d["ref_ipm_status"] and r.iterations == d["ref_ipm_iterations"]
    if d["ref_ipm_status"] == "OPTIMAL" and d["ref_ipm_iterations"] < 100:
        assert abs(r.objective_value - d["ref_ipm_objective"]) <= 1e-8 * max(1.0, abs(d["ref_ipm_objective"]))


def test_ipm_sparse_solver_bandm_carries_on(hip):
    # the reference ends bandm in NUMERICAL_ERROR at iteration 18 (a direction
    # that is not finite after clamped pivots); the factor here is the
    # reference's bit for bit, but the iterates reach it through reductions
    # summed in another order, and this solve carries on to the optimum
    d = FIX["bandm"]
    r = hip.ipm_solve(_netlib("bandm"), solver=hip.HN_NE_SPARSE)
    assert r.status == "OPTIMAL"
    assert abs(r.objective_value - d["highs"]) <= 1e-6 * max(1.0, abs(d["highs"]))


@need_ref
@pytest.mark.parametrize("args", [(3, 30, 40), (1, 8, 12), (5, 20, 30)])
def test_ipm_auto_takes_sparse_above_dense_max_m(hip, args):
    # transportation LPs (m = S + T) with dense_max_m below m: AUTO orders N,
    # factors it sparsely like the reference and ends where the reference's
    # own InteriorPointSolver ends (these N are singular: a balanced
    # transportation problem has one redundant row, so a pivot is clamped)
    lp = lpgen.transport_lp(*args)
    ref = O.ref_ipm_solve(lp, time_limit=60.0)
    ra = hip.ipm_solve(lp, dense_max_m=10)
    print(f"[case] transport {args}: reference {ref.status} {ref.iterations}, auto {ra.status} {ra.iterations} "
          f"nnz(L) {ra.sparse_l_nnz}", flush=True)
    assert ra.sparse_l_nnz > 0 and ra.status == ref.status and ra.iterations == ref.iterations
    tol = 1e-8 if ref.iterations < 100 else 1e-5
    assert abs(ra.objective_value - ref.objective) <= tol * max(1.0, abs(ref.objective))


def test_ipm_auto_falls_back_to_pcg_when_the_ordering_fills(hip):
    # the power-law LP family fills its factor almost completely; with
    # dense_max_m below m, AUTO's ordering stops at its work budget and the
    # solve runs on PCG (no sparse factor)
    lp = lpgen.powerlaw_lp(3, 3000, 7500, 2.179, 2, 200)
    r = hip.ipm_solve(lp, dense_max_m=100, max_iterations=3)
    assert r.sparse_l_nnz == 0 and r.cg_iterations > 0


def test_ipm_sparse_solver_rejects_other_modes(hip):
    lp = lpgen.random_lp(1, 50, 120)
    with pytest.raises(RuntimeError):
        hip.ipm_solve(lp, solver=hip.HN_NE_SPARSE, pivot_rule=hip.HN_PIVOT_SKIP)
    with pytest.raises(RuntimeError):
        hip.ipm_solve(lp, solver=hip.HN_NE_SPARSE, precision=hip.HN_PREC_MIXED)
