"""Exact minimum-degree ordering and symbolic LDL' (hn_sldl_analyze, host).

The reference orders the normal equations by exact minimum degree with the
lowest index breaking ties and derives L's row patterns from the elimination
(compute_md_order_and_pattern, core/numerical/factorization.cpp:71-191).
Checks, all bit-exact:
  * the oracle's literal restatement against the reference's own
    SparseCholesky compiled into oracle/_ref (perm, L row pointer, L
    columns), on the reference's unit-test matrices (test_numerical.cpp:
    479-582) and on normal-equation patterns of real LPs;
  * the library's analysis against both, on the same and larger patterns;
  * the elimination tree and postorder against the oracle's restatement of
    compute_elimination_tree / compute_postorder (:1327-1362), which the
    reference computes but does not expose;
  * the level schedule the device factor runs on: every row's entries name
    rows of lower levels only;
  * the work limit stops the analysis and reports it.
No device is needed."""
import os

import numpy as np
import pytest

import hypernova_hip as H
import lpgen
import mps
import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


def _cycle4():
    # the 4x4 cycle 0-1-3-2-0 of SparseLDLTFillInPattern (test_numerical.cpp:479-490)
    rp = np.array([0, 3, 6, 9, 12])
    ci = np.array([0, 1, 2, 0, 1, 3, 0, 2, 3, 1, 2, 3])
    return 4, rp, ci


def _grid8():
    # the 2x4 grid Laplacian of SparseCholeskyRefactorizeReuse (:530-546)
    pairs = [(0, 1), (0, 4), (1, 2), (1, 5), (2, 3), (2, 6), (3, 7), (4, 5), (5, 6), (6, 7)]
    rows = [{i} for i in range(8)]
    for i, j in pairs:
        rows[i].add(j)
        rows[j].add(i)
    return _csr(8, rows)


def _csr(n, rows):
    rp = np.zeros(n + 1, np.int64)
    ci = []
    for i in range(n):
        r = sorted(rows[i])
        ci += r
        rp[i + 1] = rp[i] + len(r)
    return n, rp, np.array(ci, np.int64)


def _random_sym(n, density, seed):
    rng = np.random.default_rng(seed)
    rows = [{i} for i in range(n)]
    for i in range(n):
        for j in range(i + 1, n):
            if rng.random() < density:
                rows[i].add(j)
                rows[j].add(i)
    return _csr(n, rows)


def _ne_pattern(lp):
    sf = O.standard_form(lp)
    rp, ci = O.adat_pattern(sf["rp"], sf["ci"], sf["m"], sf["n"])
    return sf["m"], rp, ci


def _netlib(name):
    lp, _ = mps.read_mps(os.path.join(HERE, "golden", "netlib", name + ".mps.gz"))
    return lp


SMALL = {"cycle4": _cycle4, "grid8": _grid8}
for _n in range(3, 13):
    SMALL[f"random_{_n}"] = (lambda n=_n: _random_sym(n, 0.3, n))
SMALL["random_40"] = lambda: _random_sym(40, 0.08, 40)
SMALL["afiro"] = lambda: _ne_pattern(_netlib("afiro"))
SMALL["sc50a"] = lambda: _ne_pattern(_netlib("sc50a"))
SMALL["blend"] = lambda: _ne_pattern(_netlib("blend"))
SMALL["transport_6x9"] = lambda: _ne_pattern(lpgen.transport_lp(1, 6, 9))

LARGER = {
    "sctap1": lambda: _ne_pattern(_netlib("sctap1")),
    "bandm": lambda: _ne_pattern(_netlib("bandm")),
    "share1b": lambda: _ne_pattern(_netlib("share1b")),
    "random_lp_200x500": lambda: _ne_pattern(lpgen.random_lp(3, 200, 500)),
    "transport_30x40": lambda: _ne_pattern(lpgen.transport_lp(2, 30, 40)),
}

need_ref = pytest.mark.skipif(not O.have_reference() or not hasattr(O.REF, "ref_chol_create"),
                              reason="oracle/_ref (the reference build) not present")


def _ref(n, rp, ci):
    return O.ref_sparse_cholesky(rp, ci, np.ones(len(ci)))


@need_ref
@pytest.mark.parametrize("name", sorted(SMALL))
def test_oracle_md_order_matches_reference(name):
    n, rp, ci = SMALL[name]()
    perm, lrp, lci = O.md_order(n, rp, ci)
    R = _ref(n, rp, ci)
    assert np.array_equal(perm, R["perm"])
    assert np.array_equal(lrp, R["l_row_ptr"]) and np.array_equal(lci, R["l_col"])


@pytest.mark.parametrize("name", sorted(SMALL) + sorted(LARGER))
def test_md_order_matches_oracle_and_reference(name):
    n, rp, ci = (SMALL.get(name) or LARGER[name])()
    s = H.SparseLDL(rp, ci)
    nn, lnnz, levels, aborted = s.sizes()
    assert nn == n and not aborted
    sy = s.symbolic()
    if name in SMALL:
        perm, lrp, lci = O.md_order(n, rp, ci)
        assert np.array_equal(sy["perm"], perm)
        assert np.array_equal(sy["l_row_ptr"], lrp) and np.array_equal(sy["l_col"], lci)
    if O.have_reference() and hasattr(O.REF, "ref_chol_create"):
        R = _ref(n, rp, ci)
        assert np.array_equal(sy["perm"], R["perm"])
        assert np.array_equal(sy["l_row_ptr"], R["l_row_ptr"]) and np.array_equal(sy["l_col"], R["l_col"])
    # a permutation; every row of L ends with its diagonal and is ascending
    assert np.array_equal(np.sort(sy["perm"]), np.arange(n))
    lrp, lci = sy["l_row_ptr"], sy["l_col"]
    for i in range(n):
        row = lci[lrp[i]:lrp[i + 1]]
        assert row[-1] == i and np.all(np.diff(row) > 0)
    parent, post = O.etree_postorder(lrp, lci)
    assert np.array_equal(sy["parent"], parent) and np.array_equal(sy["postorder"], post)
    print(f"[case] {name}: n={n} nnz(N)={len(ci)} nnz(L)={lnnz} levels={levels}")


def test_md_order_fill_and_reconstruction():
    # SparseLDLTFillInPattern's claim: the cycle needs symbolic fill, and the
    # pattern holds P A P' exactly (the oracle's factor reconstructs it)
    n, rp, ci = _cycle4()
    v = np.array([4.0, 1, 1, 1, 4, 1, 1, 4, 1, 1, 1, 4])
    s = H.SparseLDL(rp, ci)
    sy = s.symbolic()
    assert s.sizes()[1] > 8
    L, D, singular = O.sparse_ldl(rp, ci, v, sy["perm"], sy["l_row_ptr"], sy["l_col"])
    assert not singular
    Ld = np.zeros((n, n))
    for i in range(n):
        for k in range(sy["l_row_ptr"][i], sy["l_row_ptr"][i + 1]):
            Ld[i, sy["l_col"][k]] = L[k]
    A = np.zeros((n, n))
    for i in range(n):
        A[i, ci[rp[i]:rp[i + 1]]] = v[rp[i]:rp[i + 1]]
    P = sy["perm"]
    assert np.abs(Ld @ np.diag(D) @ Ld.T - A[np.ix_(P, P)]).max() < 1e-12


def test_md_order_levels_respect_dependencies():
    n, rp, ci = LARGER["sctap1"]()
    s = H.SparseLDL(rp, ci)
    sy = s.symbolic()
    lrp, lci = sy["l_row_ptr"], sy["l_col"]
    level = np.zeros(n, np.int64)
    for i in range(n):
        deps = lci[lrp[i]:lrp[i + 1] - 1]
        level[i] = 1 + level[deps].max() if deps.size else 0
    assert s.sizes()[2] == level.max() + 1


def test_md_order_chain_and_empty():
    # a path graph keeps its order (every vertex of least degree is the
    # lowest end), no fill; an empty pattern is the identity
    n = 50
    rows = [{i} | ({i - 1} if i else set()) | ({i + 1} if i + 1 < n else set()) for i in range(n)]
    n, rp, ci = _csr(n, rows)
    sy = H.SparseLDL(rp, ci).symbolic()
    assert np.array_equal(sy["perm"], np.arange(n)) and len(sy["l_col"]) == 2 * n - 1
    e = H.SparseLDL(np.zeros(6, np.int64), np.zeros(0, np.int64))
    assert e.sizes()[:2] == (5, 5) and np.array_equal(e.symbolic()["perm"], np.arange(5))


def test_md_order_work_limit_stops():
    # a dense-filling normal-equation pattern: the ordering stops at its limit
    n, rp, ci = _ne_pattern(lpgen.random_lp(1, 400, 1000))
    s = H.SparseLDL(rp, ci, work_limit=1e4)
    assert s.sizes()[3]
    with pytest.raises(RuntimeError):
        s.symbolic()
