"""ctypes access to the checkers (TEST INFRASTRUCTURE ONLY).

* ``liboracle.so`` - the C restatement of the reference IPM path (hn_oracle.c).
* ``_ref/libhnref.so`` - the reference's own sources compiled by this
  directory's Makefile, driven through ref_driver.cpp.

Only tests/, ``__graft_entry__.smoke()`` and the cpu_baseline leg of
bench.py import this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libhnref.so")

_i64 = C.c_int64
_i64p = C.POINTER(C.c_int64)
_u64p = C.POINTER(C.c_uint64)
_dp = C.POINTER(C.c_double)
_i8p = C.POINTER(C.c_int8)

STATUS = ["UNKNOWN", "OPTIMAL", "INFEASIBLE", "UNBOUNDED", "SUBOPTIMAL", "TIME_LIMIT",
          "ITER_LIMIT", "NUMERICAL_ERROR", "INTERRUPTED"]


def build() -> None:
    """Compiles liboracle.so and, where the reference sources exist, _ref/."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


class _Problem(C.Structure):
    _fields_ = [("m", _i64), ("n", _i64), ("rp", _i64p), ("ci", _i64p), ("v", _dp),
                ("sense", _i8p), ("rhs", _dp), ("lb", _dp), ("ub", _dp), ("obj", _dp),
                ("maximize", C.c_int), ("obj_offset", C.c_double)]


class _Options(C.Structure):
    _fields_ = [("max_iterations", C.c_int), ("time_limit_seconds", C.c_double),
                ("convergence_tol", C.c_double), ("complementarity_tol", C.c_double),
                ("regularization", C.c_double), ("interrupt", C.c_int), ("pivot_rule", C.c_int),
                ("x0", _dp), ("y0", _dp), ("s0", _dp)]


class _Result(C.Structure):
    _fields_ = [("status", C.c_int), ("iterations", _i64), ("objective", C.c_double),
                ("primal", _dp), ("sf_m", _i64), ("sf_n", _i64), ("x", _dp), ("lam", _dp),
                ("z", _dp), ("cap_m", _i64), ("cap_n", _i64), ("rp_norm", C.c_double),
                ("rd_norm", C.c_double), ("mu", C.c_double), ("trace", _dp), ("trace_cap", _i64)]


class _VerifyOut(C.Structure):
    _fields_ = [("feasible", C.c_int), ("optimal", C.c_int), ("primal_infeasibility", C.c_double),
                ("dual_infeasibility", C.c_double), ("complementarity", C.c_double),
                ("integrality_violation", C.c_double), ("objective_discrepancy", C.c_double),
                ("n_violated", _i64), ("violated", _i64p), ("violated_cap", _i64), ("message", C.c_char * 160)]


def _load_oracle():
    if not os.path.exists(ORACLE_SO):
        build()
    lib = C.CDLL(ORACLE_SO)
    lib.or_spmv.argtypes = [_i64, _i64p, _i64p, _dp, _dp, _dp]
    lib.or_spmv_t.argtypes = [_i64, _i64, _i64p, _i64p, _dp, _dp, _dp]
    lib.or_spmm.argtypes = [_i64, _i64p, _i64p, _dp, _dp, _dp, _i64]
    lib.or_csr_transpose.argtypes = [_i64, _i64, _i64p, _i64p, _dp, _i64p, _i64p, _dp]
    lib.or_dot.argtypes = [_i64, _dp, _dp]
    lib.or_dot.restype = C.c_double
    lib.or_axpy.argtypes = [_i64, C.c_double, _dp, _dp]
    lib.or_norm_inf.argtypes = [_i64, _dp]
    lib.or_norm_inf.restype = C.c_double
    lib.or_max_step.argtypes = [_i64, _dp, _dp]
    lib.or_max_step.restype = C.c_double
    lib.or_adat_pattern.argtypes = [_i64, _i64, _i64p, _i64p, _dp, _i64p, _i64p]
    lib.or_adat_pattern.restype = _i64
    lib.or_adat_dense.argtypes = [_i64, _i64, _i64p, _i64p, _dp, _dp, C.c_double, _dp]
    lib.or_adat_dense.restype = C.c_double
    lib.or_default_options.argtypes = [C.POINTER(_Options)]
    lib.or_ipm_solve.argtypes = [C.POINTER(_Problem), C.POINTER(_Options), C.POINTER(_Result)]
    lib.or_ipm_solve.restype = C.c_int
    lib.or_standard_form.argtypes = [C.POINTER(_Problem), _i64p, _i64p, _i64p, _i64p, _i64p, _dp, _dp,
                                     _dp, _dp, _dp]
    lib.or_standard_form.restype = C.c_int
    lib.or_verify.argtypes = [C.POINTER(_Problem), _i8p, _dp, _i64, _dp, _i64, _dp, _i64, C.c_double, _dp,
                              C.POINTER(_VerifyOut)]
    lib.or_verify.restype = None
    return lib


OR = _load_oracle()


def _ref_lib():
    if not os.path.exists(REF_SO):
        return None
    lib = C.CDLL(REF_SO)
    lib.ref_ipm_solve.argtypes = [_i64, _i64, _i64p, _i64p, _dp, _i8p, _dp, _dp, _dp, _dp, C.c_int,
                                  C.c_double, C.c_int, C.c_double, C.c_int, C.POINTER(C.c_int), _i64p,
                                  _dp, _dp, _dp, _i64, _i64p, _dp]
    lib.ref_cpu_spmv.argtypes = [_i64, _i64, _u64p, _u64p, _dp, _i64, _dp, _dp, C.c_int]
    lib.ref_cpu_spmv_time.argtypes = [_i64, _i64, _u64p, _u64p, _dp, _i64, _dp, C.c_int]
    lib.ref_cpu_spmv_time.restype = C.c_double
    lib.ref_cpu_spmv_t_time.argtypes = [_i64, _i64, _i64, _u64p, _u64p, _dp, _dp, C.c_int]
    lib.ref_cpu_spmv_t_time.restype = C.c_double
    lib.ref_transpose.argtypes = [_i64, _i64, _i64p, _i64p, _dp, _i64p, _i64p, _dp]
    lib.ref_transpose.restype = _i64
    if hasattr(lib, "ref_chol_create"):
        lib.ref_chol_create.argtypes = [_i64, _u64p, _u64p, _dp, _i64p, C.POINTER(C.c_int)]
        lib.ref_chol_create.restype = C.c_void_p
        lib.ref_chol_get.argtypes = [C.c_void_p, _u64p, _u64p, _u64p, _dp, _dp]
        lib.ref_chol_solve.argtypes = [C.c_void_p, _dp, _dp]
        lib.ref_chol_free.argtypes = [C.c_void_p]
        lib.ref_chol_free.restype = None
    if hasattr(lib, "ref_crossover"):
        lib.ref_ipm_solve_crossover.argtypes = [_i64, _i64, _i64p, _i64p, _dp, _i8p, _dp, _dp, _dp, _dp, C.c_int,
                                                C.c_double, C.c_int, C.c_int, C.POINTER(C.c_int), _i64p, _dp, _dp,
                                                _dp]
        lib.ref_crossover.argtypes = [_i64, _i64, _i64p, _i64p, _dp, _i8p, _dp, _dp, _dp, _dp, C.c_int, C.c_double,
                                      C.c_int, C.c_double, _i64, _dp, C.c_int, C.POINTER(C.c_int), _i64p, _dp, _dp,
                                      _dp]
    if hasattr(lib, "ref_presolve"):
        lib.ref_presolve.argtypes = [_i64, _i64, _i64p, _i64p, _dp, _i8p, _dp, _dp, _dp, _dp, C.c_int, C.c_double,
                                     C.c_int, C.POINTER(C.c_void_p), _i64p, _dp]
        lib.ref_presolve_reduced.argtypes = [C.c_void_p, _i64p, _i64p, _dp, _i8p, _dp, _dp, _dp, _dp,
                                             C.POINTER(C.c_int), _dp]
        lib.ref_postsolve.argtypes = [C.c_void_p, _dp, _i64, _dp, _i64, _dp, _dp, _dp, _dp]
        lib.ref_presolve_free.argtypes = [C.c_void_p]
        lib.ref_presolve_free.restype = None
    if hasattr(lib, "ref_verify"):
        lib.ref_verify.argtypes = [_i64, _i64, _i64p, _i64p, _dp, _i8p, _dp, _dp, _dp, _dp, C.c_int, C.c_double,
                                   _i8p, _dp, _i64, _dp, _i64, _dp, _i64, C.c_double, _dp, C.POINTER(C.c_int),
                                   _dp, _i64p, _i64, _i64p, C.c_char_p, _i64, _dp]
    return lib


REF = _ref_lib()


def have_reference() -> bool:
    return REF is not None


def _p(a, t):
    return a.ctypes.data_as(t)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i64a(a):
    return np.ascontiguousarray(a, dtype=np.int64)


# ---- kernels --------------------------------------------------------------
def spmv(rp, ci, v, x) -> np.ndarray:
    rp, ci, v, x = _i64a(rp), _i64a(ci), _f64(v), _f64(x)
    y = np.empty(rp.size - 1)
    OR.or_spmv(rp.size - 1, _p(rp, _i64p), _p(ci, _i64p), _p(v, _dp), _p(x, _dp), _p(y, _dp))
    return y


def spmv_t(rp, ci, v, x, ncols: int) -> np.ndarray:
    rp, ci, v, x = _i64a(rp), _i64a(ci), _f64(v), _f64(x)
    out = np.empty(ncols)
    OR.or_spmv_t(rp.size - 1, ncols, _p(rp, _i64p), _p(ci, _i64p), _p(v, _dp), _p(x, _dp), _p(out, _dp))
    return out


def spmm(rp, ci, v, B, k: int) -> np.ndarray:
    rp, ci, v, B = _i64a(rp), _i64a(ci), _f64(v), _f64(B).ravel()
    out = np.empty((rp.size - 1) * k)
    OR.or_spmm(rp.size - 1, _p(rp, _i64p), _p(ci, _i64p), _p(v, _dp), _p(B, _dp), _p(out, _dp), k)
    return out


def csr_transpose(rp, ci, v, ncols: int):
    rp, ci, v = _i64a(rp), _i64a(ci), _f64(v)
    orp = np.empty(ncols + 1, np.int64)
    oci = np.empty(ci.size, np.int64)
    ov = np.empty(ci.size)
    OR.or_csr_transpose(rp.size - 1, ncols, _p(rp, _i64p), _p(ci, _i64p), _p(v, _dp), _p(orp, _i64p),
                        _p(oci, _i64p), _p(ov, _dp))
    return orp, oci, ov


def dot(a, b) -> float:
    a, b = _f64(a), _f64(b)
    return OR.or_dot(min(a.size, b.size), _p(a, _dp), _p(b, _dp))


def norm_inf(v) -> float:
    v = _f64(v)
    return OR.or_norm_inf(v.size, _p(v, _dp))


def max_step(v, d) -> float:
    v, d = _f64(v), _f64(d)
    return OR.or_max_step(v.size, _p(v, _dp), _p(d, _dp))


def adat_pattern(rp, ci, m: int, n: int, D=None):
    rp, ci = _i64a(rp), _i64a(ci)
    orp = np.empty(m + 1, np.int64)
    Dp = _p(_f64(D), _dp) if D is not None else None
    nnz = OR.or_adat_pattern(m, n, _p(rp, _i64p), _p(ci, _i64p), Dp, _p(orp, _i64p), None)
    oci = np.empty(nnz, np.int64)
    OR.or_adat_pattern(m, n, _p(rp, _i64p), _p(ci, _i64p), Dp, _p(orp, _i64p), _p(oci, _i64p))
    return orp, oci


def adat_dense(rp, ci, v, D, m: int, n: int, reg: float = 1e-8):
    rp, ci, v, D = _i64a(rp), _i64a(ci), _f64(v), _f64(D)
    N = np.empty(m * m)
    dyn = OR.or_adat_dense(m, n, _p(rp, _i64p), _p(ci, _i64p), _p(v, _dp), _p(D, _dp), reg, _p(N, _dp))
    return N.reshape(m, m), dyn


# ---- sparse LDL' on the minimum-degree ordering (pure Python, small n) ------
def md_order(n: int, rp, ci):
    """compute_md_order_and_pattern (core/numerical/factorization.cpp:103-191)
    over build_elimination_graph (:71-97), restated literally: adjacency lists
    keep self loops and dead vertices, deg counts live neighbours, the next
    vertex is the live one of least degree (lowest index on ties), its live
    neighbour pairs gain edges.  Returns (perm, l_row_ptr, l_col)."""
    adj = [[] for _ in range(n)]
    for i in range(n):
        seen = set()
        for k in range(int(rp[i]), int(rp[i + 1])):
            j = int(ci[k])
            if j >= n or j in seen:
                continue
            seen.add(j)
            adj[i].append(j)
    for i in range(n):
        for j in list(adj[i]):
            if j != i:
                adj[j].append(i)
    adj = [sorted(set(a)) for a in adj]
    live = [True] * n
    deg = [sum(1 for w in adj[v] if w != v) for v in range(n)]
    perm, inv = [n] * n, [n] * n
    pattern = []
    for r in range(n):
        best, best_deg = n, n + 1
        for v in range(n):
            if live[v] and (deg[v] < best_deg or (deg[v] == best_deg and v < best)):
                best, best_deg = v, deg[v]
        perm[r], inv[best] = best, r
        pattern.append(sorted({inv[w] for w in adj[best] if w != best and inv[w] < r} | {r}))
        live[best] = False
        for w in adj[best]:
            if live[w]:
                deg[w] -= 1
        nb = list(adj[best])
        for a in nb:
            if not live[a] or a == best:
                continue
            for b in nb:
                if not live[b] or b <= a:
                    continue
                if b not in adj[a]:
                    adj[a].append(b)
                    adj[b].append(a)
                    deg[a] += 1
                    deg[b] += 1
    lrp = np.zeros(n + 1, np.int64)
    for i in range(n):
        lrp[i + 1] = lrp[i] + len(pattern[i])
    lci = np.array([c for row in pattern for c in row], np.int64)
    return np.array(perm, np.int64), lrp, lci


def etree_postorder(lrp, lci):
    """compute_elimination_tree / compute_postorder (factorization.cpp:
    1327-1362): parent[j] = least row i > j with an entry (i, j), else n; the
    postorder of the recursive depth-first search over the row patterns."""
    n = len(lrp) - 1
    parent = [n] * n
    for i in range(n):
        for k in range(int(lrp[i]), int(lrp[i + 1])):
            j = int(lci[k])
            if j < i:
                parent[j] = min(parent[j], i)
    visited, post = [False] * n, []
    import sys
    sys.setrecursionlimit(max(1000, 4 * n + 100))

    def dfs(u):
        visited[u] = True
        for k in range(int(lrp[u]), int(lrp[u + 1])):
            v = int(lci[k])
            if v < u and not visited[v]:
                dfs(v)
        post.append(u)

    for i in range(n):
        if not visited[i]:
            dfs(i)
    return np.array(parent, np.int64), np.array(post, np.int64)


def sparse_ldl(rp, ci, v, perm, lrp, lci, singular_tol: float = 1e-10):
    """SparseCholesky::numeric_factorization (factorization.cpp:1248-1325) on
    B = P A P' (permute_symmetric :196-234; duplicates summed), restated:
    returns (L values, D, singular)."""
    n = len(perm)
    inv = np.empty(n, np.int64)
    inv[np.asarray(perm)] = np.arange(n)
    rows = [[] for _ in range(n)]
    for i in range(n):
        for k in range(int(rp[i]), int(rp[i + 1])):
            j = int(ci[k])
            if j < n:
                rows[int(inv[i])].append((int(inv[j]), float(v[k])))
    B = [dict() for _ in range(n)]
    for r in range(n):
        for c, val in sorted(rows[r]):  # pairs sorted as std::sort sorts them, equal columns summed in that order
            B[r][c] = B[r][c] + val if c in B[r] else val
    L = np.zeros(int(lrp[n]))
    D = np.zeros(n)
    singular = False
    for i in range(n):
        diag = B[i].get(i, 0.0)
        for k in range(int(lrp[i]), int(lrp[i + 1])):
            j = int(lci[k])
            if j < i:
                s = B[i].get(j, 0.0)
                kj, ki = int(lrp[j]), int(lrp[i])
                while kj < int(lrp[j + 1]) and ki < k:
                    if lci[kj] == lci[ki]:
                        s -= L[kj] * L[ki] * D[int(lci[kj])]
                        kj += 1
                        ki += 1
                    elif lci[kj] < lci[ki]:
                        kj += 1
                    else:
                        ki += 1
                L[k] = s / D[j]
            elif j == i:
                s = 0.0
                for kk in range(int(lrp[i]), k):
                    s += L[kk] * L[kk] * D[int(lci[kk])]
                val = diag - s
                if val <= singular_tol:
                    val, singular = singular_tol, True
                D[i] = val
                L[k] = 1.0
    return L, D, singular


def sparse_ldl_solve(perm, lrp, lci, L, D, b):
    """SparseCholesky::solve (factorization.cpp:1364-1396), restated."""
    n = len(perm)
    x = [float(b[int(perm[i])]) for i in range(n)]
    for i in range(n):
        s = x[i]
        for k in range(int(lrp[i]), int(lrp[i + 1])):
            j = int(lci[k])
            if j < i:
                s -= L[k] * x[j]
        x[i] = s
    for i in range(n):
        if D[i] != 0.0:
            x[i] /= D[i]
    for i in range(n - 1, -1, -1):
        for k in range(int(lrp[i]), int(lrp[i + 1])):
            j = int(lci[k])
            if j < i:
                x[j] -= L[k] * x[i]
    out = np.empty(n)
    out[np.asarray(perm)] = x
    return out


# ---- IPM ------------------------------------------------------------------
@dataclass
class IpmResult:
    status: str
    iterations: int
    objective: float
    primal: np.ndarray
    x: Optional[np.ndarray] = None
    lam: Optional[np.ndarray] = None
    z: Optional[np.ndarray] = None
    rp_norm: float = 0.0
    rd_norm: float = 0.0
    mu: float = 0.0
    trace: Optional[np.ndarray] = None
    solve_ms: float = 0.0
    dual: Optional[np.ndarray] = None


def _problem(lp):
    keep = dict(rp=_i64a(lp.row_ptr), ci=_i64a(lp.col), v=_f64(lp.val), sense=np.ascontiguousarray(lp.sense, np.int8),
                rhs=_f64(lp.rhs), lb=_f64(lp.lb), ub=_f64(lp.ub), obj=_f64(lp.obj))
    p = _Problem(lp.m, lp.n, _p(keep["rp"], _i64p), _p(keep["ci"], _i64p), _p(keep["v"], _dp),
                 _p(keep["sense"], _i8p), _p(keep["rhs"], _dp), _p(keep["lb"], _dp), _p(keep["ub"], _dp),
                 _p(keep["obj"], _dp), int(lp.maximize), float(getattr(lp, "obj_offset", 0.0)))
    return p, keep


def standard_form(lp):
    """The scaled standard form (M, rhs, c, col_scale, offset) the IPM iterates on."""
    p, keep = _problem(lp)
    sm, sn, sz = _i64(), _i64(), _i64()
    OR.or_standard_form(C.byref(p), C.byref(sm), C.byref(sn), C.byref(sz), None, None, None, None, None, None,
                        None)
    rp = np.empty(sm.value + 1, np.int64)
    ci = np.empty(sz.value, np.int64)
    v = np.empty(sz.value)
    rhs = np.empty(sm.value)
    c = np.empty(sn.value)
    cs = np.empty(sn.value)
    off = C.c_double()
    OR.or_standard_form(C.byref(p), C.byref(sm), C.byref(sn), C.byref(sz), _p(rp, _i64p), _p(ci, _i64p),
                        _p(v, _dp), _p(rhs, _dp), _p(c, _dp), _p(cs, _dp), C.byref(off))
    return dict(m=sm.value, n=sn.value, rp=rp, ci=ci, v=v, rhs=rhs, c=c, col_scale=cs, offset=off.value)


def ipm_solve(lp, max_iterations: int = 100, time_limit: float = 0.0, interrupt: bool = False,
              trace_cap: int = 200, pivot_rule: int = 0, warm=None) -> IpmResult:
    """The restated reference IPM (crossover off).  pivot_rule=1 selects the
    skip rule that is not in the reference (r = reg, tiny pivots skipped);
    warm = (x0, y0, s0) of the scaled standard form replaces the start point."""
    p, keep = _problem(lp)
    o = _Options()
    OR.or_default_options(C.byref(o))
    o.max_iterations = max_iterations
    o.time_limit_seconds = time_limit
    o.interrupt = int(interrupt)
    o.pivot_rule = int(pivot_rule)
    if warm is not None:
        keep["warm"] = [_f64(a) for a in warm]
        o.x0, o.y0, o.s0 = (_p(a, _dp) for a in keep["warm"])
    cap_m = lp.m + lp.n + 8
    cap_n = 2 * lp.n + lp.m + 8
    primal = np.zeros(max(lp.n, 1))
    x = np.zeros(cap_n)
    z = np.zeros(cap_n)
    lam = np.zeros(cap_m)
    trace = np.zeros(5 * trace_cap)
    r = _Result()
    r.primal = _p(primal, _dp)
    r.x, r.z, r.lam = _p(x, _dp), _p(z, _dp), _p(lam, _dp)
    r.cap_m, r.cap_n = cap_m, cap_n
    r.trace = _p(trace, _dp)
    r.trace_cap = trace_cap
    if OR.or_ipm_solve(C.byref(p), C.byref(o), C.byref(r)) != 0:
        raise ValueError("oracle rejected the problem")
    it = int(r.iterations)
    return IpmResult(status=STATUS[r.status], iterations=it, objective=r.objective, primal=primal[:lp.n].copy(),
                     x=x[:r.sf_n].copy(), lam=lam[:r.sf_m].copy(), z=z[:r.sf_n].copy(), rp_norm=r.rp_norm,
                     rd_norm=r.rd_norm, mu=r.mu, trace=trace[:5 * min(it, trace_cap)].reshape(-1, 5).copy())


def ref_ipm_solve(lp, max_iterations: int = 100, time_limit: float = 0.0, interrupt: bool = False,
                  trace_cap: int = 200) -> IpmResult:
    """The reference's own InteriorPointSolver (crossover off), from _ref/."""
    if REF is None:
        raise RuntimeError("reference build not present")
    keep = dict(rp=_i64a(lp.row_ptr), ci=_i64a(lp.col), v=_f64(lp.val), sense=np.ascontiguousarray(lp.sense, np.int8),
                rhs=_f64(lp.rhs), lb=_f64(lp.lb), ub=_f64(lp.ub), obj=_f64(lp.obj))
    st = C.c_int()
    it = _i64()
    obj = C.c_double()
    primal = np.zeros(max(lp.n, 1))
    trace = np.zeros(5 * trace_cap)
    tl = _i64()
    ms = C.c_double()
    # trace_cap 0: no trace, which also leaves the process's stderr and
    # environment alone, so several threads may solve at once
    REF.ref_ipm_solve(lp.m, lp.n, _p(keep["rp"], _i64p), _p(keep["ci"], _i64p), _p(keep["v"], _dp),
                      _p(keep["sense"], _i8p), _p(keep["rhs"], _dp), _p(keep["lb"], _dp), _p(keep["ub"], _dp),
                      _p(keep["obj"], _dp), int(lp.maximize), float(getattr(lp, "obj_offset", 0.0)),
                      max_iterations, time_limit, int(interrupt), C.byref(st), C.byref(it), C.byref(obj),
                      _p(primal, _dp), _p(trace, _dp) if trace_cap > 0 else None, trace_cap, C.byref(tl),
                      C.byref(ms))
    return IpmResult(status=STATUS[st.value], iterations=int(it.value), objective=obj.value,
                     primal=primal[:lp.n].copy(), trace=trace[:5 * tl.value].reshape(-1, 5).copy(),
                     solve_ms=ms.value)


def ref_cpu_spmv(rp, ci, v, x, repeats: int = 1) -> np.ndarray:
    if REF is None:
        raise RuntimeError("reference build not present")
    rp = np.ascontiguousarray(rp, np.uint64)
    ci = np.ascontiguousarray(ci, np.uint64)
    v, x = _f64(v), _f64(x)
    y = np.empty(rp.size - 1)
    REF.ref_cpu_spmv(rp.size - 1, v.size, _p(rp, _u64p), _p(ci, _u64p), _p(v, _dp), x.size, _p(x, _dp),
                     _p(y, _dp), repeats)
    return y


def ref_cpu_spmv_seconds(rp, ci, v, x, repeats: int) -> float:
    """Wall seconds of `repeats` reference CPU spmv calls (1 thread)."""
    rp = np.ascontiguousarray(rp, np.uint64)
    ci = np.ascontiguousarray(ci, np.uint64)
    v, x = _f64(v), _f64(x)
    return REF.ref_cpu_spmv_time(rp.size - 1, v.size, _p(rp, _u64p), _p(ci, _u64p), _p(v, _dp), x.size,
                                 _p(x, _dp), repeats)


def ref_cpu_spmv_t_seconds(rp, ci, v, w, ncols: int, repeats: int) -> float:
    """Wall seconds of `repeats` reference transpose_multiply calls (1 thread)."""
    rp = np.ascontiguousarray(rp, np.uint64)
    ci = np.ascontiguousarray(ci, np.uint64)
    v, w = _f64(v), _f64(w)
    return REF.ref_cpu_spmv_t_time(rp.size - 1, ncols, v.size, _p(rp, _u64p), _p(ci, _u64p), _p(v, _dp),
                                   _p(w, _dp), repeats)


def port_spmv_seconds(rp, ci, v, x, repeats: int) -> float:
    """Wall seconds of `repeats` restated CPU spmv calls (1 thread)."""
    import time
    rp, ci, v, x = _i64a(rp), _i64a(ci), _f64(v), _f64(x)
    y = np.empty(rp.size - 1)
    t = time.perf_counter()
    for _ in range(repeats):
        OR.or_spmv(rp.size - 1, _p(rp, _i64p), _p(ci, _i64p), _p(v, _dp), _p(x, _dp), _p(y, _dp))
    return time.perf_counter() - t


def port_spmv_mt_seconds(rp, ci, v, x, repeats: int, threads: int) -> float:
    """Wall seconds of `repeats` products on `threads` host threads (this
    repository's row-split loop, not the reference's kernel)."""
    rp, ci, v, x = _i64a(rp), _i64a(ci), _f64(v), _f64(x)
    y = np.empty(rp.size - 1)
    OR.or_spmv_mt_seconds.argtypes = [_i64, _i64p, _i64p, _dp, _dp, _dp, C.c_int, C.c_int]
    OR.or_spmv_mt_seconds.restype = C.c_double
    return OR.or_spmv_mt_seconds(rp.size - 1, _p(rp, _i64p), _p(ci, _i64p), _p(v, _dp), _p(x, _dp), _p(y, _dp),
                                 repeats, threads)


def ref_transpose(rp, ci, v, ncols: int):
    rp, ci, v = _i64a(rp), _i64a(ci), _f64(v)
    nnz = REF.ref_transpose(rp.size - 1, ncols, _p(rp, _i64p), _p(ci, _i64p), _p(v, _dp), None, None, None)
    orp = np.empty(ncols + 1, np.int64)
    oci = np.empty(nnz, np.int64)
    ov = np.empty(nnz)
    REF.ref_transpose(rp.size - 1, ncols, _p(rp, _i64p), _p(ci, _i64p), _p(v, _dp), _p(orp, _i64p),
                      _p(oci, _i64p), _p(ov, _dp))
    return orp, oci, ov


def ref_sparse_cholesky(rp, ci, v, b=None) -> dict:
    """The reference's SparseCholesky (analyze + factorize, industrial
    tolerances) on the n x n CSR, from _ref/: perm, l_row_ptr, l_col, L, D,
    singular and, when b is given, x = A^{-1} b by its solve."""
    if REF is None or not hasattr(REF, "ref_chol_create"):
        raise RuntimeError("reference build not present")
    rp = np.ascontiguousarray(rp, np.uint64)
    ci = np.ascontiguousarray(ci, np.uint64)
    v = _f64(v)
    n = rp.size - 1
    nnz, sing = _i64(), C.c_int()
    h = REF.ref_chol_create(n, _p(rp, _u64p), _p(ci, _u64p), _p(v, _dp), C.byref(nnz), C.byref(sing))
    try:
        out = dict(perm=np.zeros(n, np.uint64), l_row_ptr=np.zeros(n + 1, np.uint64),
                   l_col=np.zeros(max(nnz.value, 1), np.uint64), L=np.zeros(max(nnz.value, 1)), D=np.zeros(max(n, 1)))
        REF.ref_chol_get(h, _p(out["perm"], _u64p), _p(out["l_row_ptr"], _u64p), _p(out["l_col"], _u64p),
                         _p(out["L"], _dp), _p(out["D"], _dp))
        if b is not None:
            bb = _f64(b)
            x = np.empty(n)
            REF.ref_chol_solve(h, _p(bb, _dp), _p(x, _dp))
            out["x"] = x
    finally:
        REF.ref_chol_free(h)
    out["l_col"], out["L"], out["D"] = out["l_col"][:nnz.value], out["L"][:nnz.value], out["D"][:n]
    for k in ("perm", "l_row_ptr", "l_col"):
        out[k] = out[k].astype(np.int64)
    out["singular"] = bool(sing.value)
    return out


# ---- post-solve verification ------------------------------------------------
@dataclass
class VerifyResult:
    """validation::VerificationResult (solution_verifier.hpp:12-21)."""
    feasible: bool
    optimal: bool
    primal_infeasibility: float
    dual_infeasibility: float
    complementarity: float
    integrality_violation: float
    objective_discrepancy: float
    violated: list
    n_violated: int
    message: str
    seconds: float = 0.0


DEFAULT_TOL = (1e-9, 1e-9, 1e-6, 1e-6)  # feasibility, optimality, complementarity, integrality


def _vec(a):
    return None if a is None else _f64(a)


def verify(lp, primal, dual=None, reduced_costs=None, objective_value: float = 0.0, var_type=None,
           tol=DEFAULT_TOL, cap: int = 1 << 20) -> VerifyResult:
    """The restated SolutionVerifier::verify_detailed.  dual / reduced_costs
    of None are empty vectors, as in a default model::Solution."""
    p, keep = _problem(lp)
    primal, dual, rc = _f64(primal), _vec(dual), _vec(reduced_costs)
    vt = None if var_type is None else np.ascontiguousarray(var_type, np.int8)
    t = _f64(tol)
    viol = np.zeros(max(cap, 1), np.int64)
    out = _VerifyOut()
    out.violated = _p(viol, _i64p)
    out.violated_cap = cap
    OR.or_verify(C.byref(p), None if vt is None else _p(vt, _i8p), _p(primal, _dp), primal.size,
                 None if dual is None else _p(dual, _dp), 0 if dual is None else dual.size,
                 None if rc is None else _p(rc, _dp), 0 if rc is None else rc.size, float(objective_value),
                 _p(t, _dp), C.byref(out))
    k = min(out.n_violated, cap)
    return VerifyResult(bool(out.feasible), bool(out.optimal), out.primal_infeasibility, out.dual_infeasibility,
                        out.complementarity, out.integrality_violation, out.objective_discrepancy,
                        viol[:k].tolist(), int(out.n_violated), out.message.decode())


def ref_verify(lp, primal, dual=None, reduced_costs=None, objective_value: float = 0.0, var_type=None,
               tol=DEFAULT_TOL, cap: int = 1 << 20) -> VerifyResult:
    """The reference's own SolutionVerifier::verify_detailed, from _ref/.  The
    problem goes through ProblemBuilder::build (rows sorted, duplicates
    summed), so pass rows in that form to compare with verify()."""
    if REF is None or not hasattr(REF, "ref_verify"):
        raise RuntimeError("reference build not present")
    keep = dict(rp=_i64a(lp.row_ptr), ci=_i64a(lp.col), v=_f64(lp.val), sense=np.ascontiguousarray(lp.sense, np.int8),
                rhs=_f64(lp.rhs), lb=_f64(lp.lb), ub=_f64(lp.ub), obj=_f64(lp.obj))
    primal, dual, rc = _f64(primal), _vec(dual), _vec(reduced_costs)
    vt = None if var_type is None else np.ascontiguousarray(var_type, np.int8)
    t = _f64(tol)
    flags = (C.c_int * 2)()
    vals = np.zeros(5)
    viol = np.zeros(max(cap, 1), np.int64)
    nv = _i64()
    msg = C.create_string_buffer(512)
    sec = C.c_double()
    REF.ref_verify(lp.m, lp.n, _p(keep["rp"], _i64p), _p(keep["ci"], _i64p), _p(keep["v"], _dp),
                   _p(keep["sense"], _i8p), _p(keep["rhs"], _dp), _p(keep["lb"], _dp), _p(keep["ub"], _dp),
                   _p(keep["obj"], _dp), int(lp.maximize), float(getattr(lp, "obj_offset", 0.0)),
                   None if vt is None else _p(vt, _i8p), _p(primal, _dp), primal.size,
                   None if dual is None else _p(dual, _dp), 0 if dual is None else dual.size,
                   None if rc is None else _p(rc, _dp), 0 if rc is None else rc.size, float(objective_value),
                   _p(t, _dp), flags, _p(vals, _dp), _p(viol, _i64p), cap, C.byref(nv), msg, 512, C.byref(sec))
    k = min(nv.value, cap)
    return VerifyResult(bool(flags[0]), bool(flags[1]), *vals.tolist(), viol[:k].tolist(), int(nv.value),
                        msg.value.decode(), sec.value)


# ---- crossover (the reference's simplex polish) -------------------------------
def _lp_keep(lp):
    return dict(rp=_i64a(lp.row_ptr), ci=_i64a(lp.col), v=_f64(lp.val), sense=np.ascontiguousarray(lp.sense, np.int8),
                rhs=_f64(lp.rhs), lb=_f64(lp.lb), ub=_f64(lp.ub), obj=_f64(lp.obj))


def _lp_args(lp, keep):
    return (lp.m, lp.n, _p(keep["rp"], _i64p), _p(keep["ci"], _i64p), _p(keep["v"], _dp), _p(keep["sense"], _i8p),
            _p(keep["rhs"], _dp), _p(keep["lb"], _dp), _p(keep["ub"], _dp), _p(keep["obj"], _dp), int(lp.maximize),
            float(getattr(lp, "obj_offset", 0.0)))


def ref_ipm_solve_crossover(lp, max_iterations: int = 100, crossover_max_iter: int = 100) -> IpmResult:
    """The reference's InteriorPointSolver with crossover on (its default)."""
    if REF is None or not hasattr(REF, "ref_crossover"):
        raise RuntimeError("reference build not present")
    keep = _lp_keep(lp)
    st, it, obj = C.c_int(), _i64(), C.c_double()
    primal, dual = np.zeros(max(lp.n, 1)), np.zeros(max(lp.m, 1))
    REF.ref_ipm_solve_crossover(*_lp_args(lp, keep), max_iterations, crossover_max_iter, C.byref(st), C.byref(it),
                                C.byref(obj), _p(primal, _dp), _p(dual, _dp))
    return IpmResult(status=STATUS[st.value], iterations=int(it.value), objective=obj.value,
                     primal=primal[:lp.n].copy(), dual=dual[:lp.m].copy())


def ref_crossover(lp, status: str, objective: float, iterations: int, primal,
                  crossover_max_iter: int = 100) -> IpmResult:
    """The crossover block of the reference's solve (interior_point.cpp:786-878)
    from a given barrier endpoint, on the reference's SimplexSolver."""
    if REF is None or not hasattr(REF, "ref_crossover"):
        raise RuntimeError("reference build not present")
    keep = _lp_keep(lp)
    x = _f64(primal)
    st, it, obj = C.c_int(), _i64(), C.c_double()
    out_x, out_y = np.zeros(max(lp.n, 1)), np.zeros(max(lp.m, 1))
    rc = REF.ref_crossover(*_lp_args(lp, keep), STATUS.index(status), float(objective), int(iterations),
                           _p(x, _dp), crossover_max_iter, C.byref(st), C.byref(it), C.byref(obj), _p(out_x, _dp),
                           _p(out_y, _dp))
    if rc == -2:
        raise ValueError("problem has duplicate rows (the reference polishes the deduplicated problem)")
    return IpmResult(status=STATUS[st.value], iterations=int(it.value), objective=obj.value,
                     primal=out_x[:lp.n].copy(), dual=out_y[:lp.m].copy())


# ---- presolve / postsolve (the reference's Presolver) ----------------------------
class RefPresolve:
    """presolve::Presolver run as Solver::solve runs it (level 2 = AGGRESSIVE,
    the default).  .reduced is the reduced LP, .ms the presolve time;
    postsolve() maps a reduced (primal, dual) back and recomputes the
    objective on the original problem."""

    def __init__(self, lp, level: int = 2):
        if REF is None or not hasattr(REF, "ref_presolve"):
            raise RuntimeError("reference build not present")
        import lpgen
        keep = _lp_keep(lp)
        self._h = C.c_void_p()
        dims = np.zeros(3, np.int64)
        ms = C.c_double()
        rc = REF.ref_presolve(*_lp_args(lp, keep), level, C.byref(self._h), _p(dims, _i64p), C.byref(ms))
        if rc != 0:
            raise ValueError(f"reference presolve returned {rc}")
        self.ms = ms.value
        m, n, z = (int(d) for d in dims)
        rp, ci, v = np.zeros(m + 1, np.int64), np.zeros(max(z, 1), np.int64), np.zeros(max(z, 1))
        sense, rhs = np.zeros(max(m, 1), np.int8), np.zeros(max(m, 1))
        lb, ub, obj = np.zeros(max(n, 1)), np.zeros(max(n, 1)), np.zeros(max(n, 1))
        mx, off = C.c_int(), C.c_double()
        REF.ref_presolve_reduced(self._h, _p(rp, _i64p), _p(ci, _i64p), _p(v, _dp), _p(sense, _i8p), _p(rhs, _dp),
                                 _p(lb, _dp), _p(ub, _dp), _p(obj, _dp), C.byref(mx), C.byref(off))
        self.reduced = lpgen.LP(m=m, n=n, row_ptr=rp, col=ci[:z].copy(), val=v[:z].copy(), rhs=rhs[:m].copy(),
                                sense=sense[:m].copy(), lb=lb[:n].copy(), ub=ub[:n].copy(), obj=obj[:n].copy(),
                                maximize=bool(mx.value), obj_offset=off.value)
        self.m0, self.n0 = lp.m, lp.n

    def postsolve(self, primal, dual=None):
        x = _f64(primal)
        y = _f64(dual if dual is not None else np.zeros(self.reduced.m))
        px, py = np.zeros(max(self.n0, 1)), np.zeros(max(self.m0, 1))
        obj, ms = C.c_double(), C.c_double()
        REF.ref_postsolve(self._h, _p(x, _dp), x.size, _p(y, _dp), y.size, _p(px, _dp), _p(py, _dp), C.byref(obj),
                          C.byref(ms))
        return px[:self.n0].copy(), py[:self.m0].copy(), obj.value, ms.value

    def __del__(self):
        if getattr(self, "_h", None) and REF is not None:
            REF.ref_presolve_free(self._h)
            self._h = None


def ref_lp_parse(text: str, keep_names: bool = True) -> dict:
    """model::LPParser::parse_string of the reference (lp_parser.cpp:25-111):
    {"success", "errors", "lp" (lpgen.LP or None), "var_names", "con_names",
    "var_type"}.  Test-side only."""
    if REF is None or not hasattr(REF, "ref_lp_parse"):
        raise RuntimeError("reference build not present")
    import lpgen
    R = REF
    R.ref_lp_parse.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_void_p), _i64p]
    R.ref_lp_problem.argtypes = [C.c_void_p, _i64p, _i64p, _dp, _i8p, _dp, _dp, _dp, _dp, _i8p,
                                 C.POINTER(C.c_int), C.c_char_p, _i64]
    R.ref_lp_errors.argtypes = [C.c_void_p, C.c_char_p, _i64]
    R.ref_lp_write.argtypes = [C.c_void_p, C.c_char_p]
    R.ref_lp_free.argtypes = [C.c_void_p]
    h = C.c_void_p()
    dims = np.zeros(5, np.int64)
    data = text if isinstance(text, bytes) else text.encode()
    rc = R.ref_lp_parse(data, 1 if keep_names else 0, C.byref(h), _p(dims, _i64p))
    if rc != 0:
        raise ValueError(f"reference LP parse returned {rc}")
    try:
        m, n, z, ne, ok = (int(d) for d in dims)
        buf = C.create_string_buffer(1 << 20)
        R.ref_lp_errors(h, buf, len(buf))
        enc = "latin-1" if isinstance(text, bytes) else "utf-8"
        out = {"success": bool(ok), "errors": buf.value.decode(enc).splitlines(), "lp": None}
        if ok:
            rp, ci, v = np.zeros(m + 1, np.int64), np.zeros(max(z, 1), np.int64), np.zeros(max(z, 1))
            sense, rhs = np.zeros(max(m, 1), np.int8), np.zeros(max(m, 1))
            lb, ub, obj = np.zeros(max(n, 1)), np.zeros(max(n, 1)), np.zeros(max(n, 1))
            vt = np.zeros(max(n, 1), np.int8)
            mx = C.c_int()
            rc = R.ref_lp_problem(h, _p(rp, _i64p), _p(ci, _i64p), _p(v, _dp), _p(sense, _i8p), _p(rhs, _dp),
                                  _p(lb, _dp), _p(ub, _dp), _p(obj, _dp), _p(vt, _i8p), C.byref(mx), buf, len(buf))
            if rc != 0:
                raise ValueError(f"reference LP problem returned {rc}")
            names = buf.value.decode(enc).split("\n")
            out["lp"] = lpgen.LP(m=m, n=n, row_ptr=rp, col=ci[:z].copy(), val=v[:z].copy(), rhs=rhs[:m].copy(),
                                 sense=sense[:m].copy(), lb=lb[:n].copy(), ub=ub[:n].copy(), obj=obj[:n].copy(),
                                 maximize=bool(mx.value))
            out["var_names"], out["con_names"] = names[:n], names[n:n + m]
            out["var_type"] = vt[:n].copy()
        return out
    finally:
        R.ref_lp_free(h)


def ref_write_lp(text: str, path: str) -> None:
    """The reference's write_lp (lp_parser.cpp:603-714) of the problem its
    LPParser reads from text.  Test-side only."""
    ref_lp_parse(text)  # sets the argtypes and checks the build
    h = C.c_void_p()
    dims = np.zeros(5, np.int64)
    if REF.ref_lp_parse(text.encode(), 1, C.byref(h), _p(dims, _i64p)) != 0 or not dims[4]:
        raise ValueError("reference LP parse failed")
    try:
        if REF.ref_lp_write(h, path.encode()) != 0:
            raise OSError("reference write_lp failed")
    finally:
        REF.ref_lp_free(h)
