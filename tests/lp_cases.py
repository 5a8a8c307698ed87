"""LPs of the reference's own IPM tests (tests/unit/test_lp.cpp and
tests/integration/test_integration.cpp), built in the reference's Problem
shape (CSR rows, senses, bounds, objective)."""
import numpy as np

import lpgen

INF = np.inf
LE, GE, EQ = lpgen.LE, lpgen.GE, lpgen.EQ


def from_rows(nvars, rows, senses, rhs, lb, ub, obj, maximize=False):
    rp, ci, v = [0], [], []
    for r in rows:
        for j, a in r:
            ci.append(j)
            v.append(a)
        rp.append(len(ci))
    return lpgen.LP(m=len(rows), n=nvars, row_ptr=np.array(rp, np.int64), col=np.array(ci, np.int64),
                    val=np.array(v, float), rhs=np.array(rhs, float), sense=np.array(senses, np.int8),
                    lb=np.array(lb, float), ub=np.array(ub, float), obj=np.array(obj, float), maximize=maximize)


def simple_lp():
    """test_lp.cpp:244-259 (also :287-400): max x1 + 2 x2, x1 + x2 <= 10,
    x1 + 0 x2 <= 5, 0 <= x <= 10.  Optimum 20 at (0, 10)."""
    return from_rows(2, [[(0, 1.0), (1, 1.0)], [(0, 1.0), (1, 0.0)]], [LE, LE], [10.0, 5.0], [0, 0], [10, 10],
                     [1.0, 2.0], maximize=True)


def larger_lp():
    """test_lp.cpp:261-285: 50 vars in [0,100], 20 identical rows
    sum_{j%5==0} x_j <= 10, max sum x."""
    rows = [[(j, 1.0) for j in range(0, 50, 5)] for _ in range(20)]
    return from_rows(50, rows, [LE] * 20, [10.0] * 20, [0] * 50, [100] * 50, [1.0] * 50, maximize=True)


def nonfinite_guard_lp():
    """test_lp.cpp:402-434: badly scaled coefficients."""
    n = 12
    rows = [[(j, 1e6 * (j + 1)) for j in range(n)] for _ in range(6)]
    return from_rows(n, rows, [LE] * 6, [1e7 * (i + 1) for i in range(6)], [0] * n, [1e8] * n,
                     [1e11 * (i + 1) for i in range(n)], maximize=True)


def time_limit_lp():
    """test_lp.cpp:604-634."""
    rows = [[(j, float((i + 1) * (j + 1))) for j in range(20)] for i in range(10)]
    return from_rows(20, rows, [LE] * 10, [100.0 * (i + 1) for i in range(10)], [0] * 20, [1e4] * 20,
                     [float(i + 1) for i in range(20)], maximize=True)


def integration_lp():
    """test_integration.cpp:56-71: max 3x1 + 5x2, x1 + x2 <= 10, 2x1 + x2 <= 14,
    x in [0,10].  Optimum 50 (x2 = 10)."""
    return from_rows(2, [[(0, 1.0), (1, 1.0)], [(0, 2.0), (1, 1.0)]], [LE, LE], [10.0, 14.0], [0, 0], [10, 10],
                     [3.0, 5.0], maximize=True)


def gpu_parity_lp():
    """test_gpu_backend.cpp:349-385: max 5x0+4x1+3x2, 2x0+x1+x2<=80, x0+2x1+x2<=90, x in [0,50]."""
    return from_rows(3, [[(0, 2.0), (1, 1.0), (2, 1.0)], [(0, 1.0), (1, 2.0), (2, 1.0)]], [LE, LE], [80.0, 90.0],
                     [0, 0, 0], [50, 50, 50], [5.0, 4.0, 3.0], maximize=True)


def auto_mode_lp():
    """test_gpu_backend.cpp:305-321: max 2x0 + 3x1, x0 + x1 <= 15, x in [0,10]. Optimum 40."""
    return from_rows(2, [[(0, 1.0), (1, 1.0)]], [LE], [15.0], [0, 0], [10, 10], [2.0, 3.0], maximize=True)


def forced_mode_lp():
    """test_gpu_backend.cpp:323-339: max 3x0 + 4x1, x0 + 2x1 <= 12, x in [0,10]. Optimum 34."""
    return from_rows(2, [[(0, 1.0), (1, 2.0)]], [LE], [12.0], [0, 0], [10, 10], [3.0, 4.0], maximize=True)


def mixed_bounds_lp():
    """Covers every standard-form branch (interior_point.cpp:220-255,316-341):
    a fixed variable, lb-only, ub-only, free, boxed; LE, GE and EQ rows; a
    duplicated row (removed by remove_duplicate_rows :82-132)."""
    rows = [
        [(0, 1.0), (1, 2.0), (2, -1.0), (3, 1.0)],
        [(1, 1.0), (3, 1.0), (4, 3.0)],
        [(0, 2.0), (2, 1.0), (4, 1.0)],
        [(1, 1.0), (3, 1.0), (4, 3.0)],
        [(2, 1.0), (3, -1.0)],
    ]
    return from_rows(5, rows, [LE, GE, EQ, GE, LE], [8.0, 2.0, 5.0, 2.0, 4.0],
                     [1.0, 0.0, -INF, -INF, 0.0], [1.0, INF, 3.0, INF, 4.0], [1.0, -1.0, 2.0, 0.5, -1.0])


REFERENCE_CASES = {
    "simple_lp": simple_lp,
    "larger_lp": larger_lp,
    "nonfinite_guard": nonfinite_guard_lp,
    "time_limit_lp": time_limit_lp,
    "integration_lp": integration_lp,
    "gpu_parity_lp": gpu_parity_lp,
    "auto_mode_lp": auto_mode_lp,
    "forced_mode_lp": forced_mode_lp,
    "mixed_bounds_lp": mixed_bounds_lp,
}


def crossover_cases():
    """(name, lp, max_iterations, crossover_max_iter) for the crossover tests:
    the reference's three crossover tests (test_lp.cpp:327-400: stalled
    barrier, converged barrier, warm polish capped at one pivot) and small
    seeded LPs the reference's simplex polishes in well under a second.
    None has duplicate rows, so the polished problem is the problem itself."""
    import lpgen
    out = [("resolves_stall", simple_lp(), 1, 100), ("exact_certificate", simple_lp(), 100, 100),
           ("cold_fallback", simple_lp(), 1, 1), ("integration", integration_lp(), 100, 100)]
    for sd in (1, 2, 3):
        out.append((f"random{sd}_20x50", lpgen.random_lp(sd, 20, 50), 100, 100))
        out.append((f"random{sd}_60x150", lpgen.random_lp(sd, 60, 150), 100, 100))
    out.append(("transport5x8", lpgen.transport_lp(1000000, 5, 8), 100, 100))
    out.append(("transport10x15", lpgen.transport_lp(1000001, 10, 15), 100, 100))
    return out


def general_bounds_lp(seed=0, m=24, n=60, free=False, maximize=False):
    """A feasible, bounded LP with every column and row kind of the standard
    form: per column j by j % 5 lb-only, ub-only, boxed, free (or, with
    free=False, lb-only again: the reference's IPM stalls on split free
    columns), fixed; rows LE, GE, EQ in turn, plus one duplicated row; an
    objective offset; maximize=True states the same LP as max -c'x.
    Feasible at an interior x*, bounded because c is built from a dual point
    that meets every sign condition (y <= 0 on LE rows, >= 0 on GE rows,
    positive lower/upper multipliers, zero reduced cost on free columns)."""
    rng = np.random.default_rng(seed)
    A = np.where(rng.random((m, n)) < 0.25, rng.standard_normal((m, n)), 0.0)
    for i in range(m):
        A[i, rng.integers(0, n)] = 1.0 + rng.random()
    A[:, 3::5] += np.where(rng.random((m, len(range(3, n, 5)))) < 0.3, 1.0, 0.0)  # free columns not empty
    xs = rng.uniform(-2, 2, n)
    kind = np.arange(n) % 5
    if not free:
        kind[kind == 3] = 0
    lb = np.where((kind == 0) | (kind == 2), xs - rng.uniform(0.5, 1.5, n), -INF)
    ub = np.where((kind == 1) | (kind == 2), xs + rng.uniform(0.5, 1.5, n), INF)
    lb[kind == 4] = ub[kind == 4] = xs[kind == 4]
    sense = np.array([(LE, GE, EQ)[i % 3] for i in range(m)], np.int8)
    ax = A @ xs
    rhs = np.where(sense == LE, ax + rng.uniform(0.5, 2, m), np.where(sense == GE, ax - rng.uniform(0.5, 2, m), ax))
    y = rng.uniform(0.2, 2, m) * np.where(sense == LE, -1.0, 1.0)
    y[sense == EQ] = rng.standard_normal(int((sense == EQ).sum()))
    lower = np.where((kind == 0) | (kind == 2), rng.uniform(0.1, 2, n), 0.0)
    upper = np.where((kind == 1), rng.uniform(0.1, 2, n), 0.0)
    c = A.T @ y + lower - upper
    c[kind == 4] = rng.standard_normal(int((kind == 4).sum()))
    rows = [[(j, A[i, j]) for j in range(n) if A[i, j] != 0.0] for i in range(m)]
    rows.append(list(rows[1]))  # an exact duplicate of a GE row
    lp = from_rows(n, rows, list(sense) + [sense[1]], list(rhs) + [rhs[1]], lb, ub, -c if maximize else c,
                   maximize=maximize)
    lp.obj_offset = 1.5
    return lp
