"""Solver-level choices (api/cpp/api.cpp:110-117, :143-168, :529-547):
boundary cases of the AUTO engine dispatch, chosen so that a wrong
comparison (< for <=, a swapped threshold) fails, the backend for each
option combination, and every backend_used string; plus where the
workloads of this build land under the reference's dispatch."""
import pytest

import lpgen
import solver as S

E = S.EngineType


@pytest.mark.parametrize("n,m,nnz,want", [
    (499, 1000, 1000, E.PRIMAL_SIMPLEX),                # n < 500
    (500, 1000, 1000, E.PRIMAL_SIMPLEX),                # n = 500, sparse, but n < 10,000
    (20000, 10, 40000, E.PRIMAL_SIMPLEX),               # density exactly 0.20 is not above it ...
    (20000, 10, 40001, E.PRIMAL_SIMPLEX),               # ... and just above is simplex too
    (10000, 1000, 10000 * 1000 * 5 // 100 - 1, E.INTERIOR_POINT),  # density just below 0.05
    (10000, 1000, 10000 * 1000 * 5 // 100, E.PRIMAL_SIMPLEX),      # density exactly 0.05
    (9999, 1000, 1000, E.PRIMAL_SIMPLEX),               # one variable short of 10,000
    (10000, 1000, 1000, E.INTERIOR_POINT),
    (10000, 0, 0, E.INTERIOR_POINT),                    # no rows: density 0
])
def test_auto_dispatch_boundaries(n, m, nnz, want):
    assert S.determine_lp_engine(n, m, nnz) == want


def test_density_exactly_point_two_with_small_n():
    # 0.20 exactly with n >= 10,000 but density above 0.05: simplex
    assert S.determine_lp_engine(10000, 5, 10000) == E.PRIMAL_SIMPLEX


def test_explicit_engine_wins():
    for e in (E.PRIMAL_SIMPLEX, E.DUAL_SIMPLEX, E.INTERIOR_POINT):
        assert S.determine_lp_engine(10, 10, 10, e) == e
        assert S.determine_lp_engine(100000, 10, 10, e) == e


def test_backend_type():
    T, B = S.ComputeTarget, S.BackendType
    assert S.backend_type(S.SolverOptions(use_gpu=False, compute_target=T.CPU_GPU_FORCE)) == B.CPU
    assert S.backend_type(S.SolverOptions(compute_target=T.CPU_ONLY)) == B.CPU
    assert S.backend_type(S.SolverOptions(compute_target=T.CPU_GPU_FORCE)) == B.HIP
    assert S.backend_type(S.SolverOptions()) == B.AUTO


def test_backend_used_strings():
    T = S.ComputeTarget
    off = S.SolverOptions(use_gpu=False)
    assert S.backend_used(off, True, True) == "cpu (GPU disabled)"
    assert S.backend_used(S.SolverOptions(compute_target=T.CPU_ONLY), True, True) == "cpu (GPU disabled)"
    auto = S.SolverOptions()
    assert S.backend_used(auto, True, True) == "hip (gfx950 kernels executed)"
    assert S.backend_used(auto, False, True) == "cpu (cost model selected CPU for matrix size)"
    assert S.backend_used(auto, False, False) == "cpu (hip unavailable)"
    force = S.SolverOptions(compute_target=T.CPU_GPU_FORCE)
    assert S.backend_used(force, True, False) == "hip (gfx950 kernels executed)"
    assert S.backend_used(force, False, True) == "hip (forced)"
    assert S.backend_used(force, False, False) == "cpu (hip forced but unavailable - fallback)"


def test_where_the_workloads_land():
    """Under the reference's AUTO dispatch the 200x300 transportation LP
    (60,000 columns) and the 200,000 x 500,000 LP reach the interior-point
    engine; the 2,000 x 5,000 LP and the batch's LPs (at most 5,000 columns)
    go to the simplex unless the caller asks for the interior-point engine."""
    t = lpgen.transport_lp(1_000_000, 200, 300)
    assert S.determine_lp_engine(t.n, t.m, t.nnz) == E.INTERIOR_POINT
    assert S.determine_lp_engine(500_000, 200_000, 3_000_000) == E.INTERIOR_POINT
    a = lpgen.random_lp(1, 2000, 5000)
    assert S.determine_lp_engine(a.n, a.m, a.nnz) == E.PRIMAL_SIMPLEX
    assert S.determine_lp_engine(5000, 2000, 25000) == E.PRIMAL_SIMPLEX
