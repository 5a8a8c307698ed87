"""GPU interior-point solver (hn_ipm_solve) against the oracle.

Bar (north star): same status, objective within 1e-8 relative, and the
primal/dual residuals and gap <= 1e-8, on the same inputs.  The iteration
count is compared too: the Newton steps agree to rounding, so the device
takes the oracle's number of steps (within one on well-conditioned LPs).
"""
import json
import os

import numpy as np
import pytest

import lp_cases
import lpgen
import oracle as O

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REF_IPM = json.load(open(os.path.join(GOLDEN, "ipm_reference.json")))

WELL = ["simple_lp", "integration_lp", "gpu_parity_lp", "auto_mode_lp", "forced_mode_lp", "random_60x150_s1",
        "random_60x150_s2", "random_60x150_s3", "random_60x150_s4", "random_60x150_s5", "random_250x625_s7",
        "powerlaw_120x300_s3"]
HARD = ["larger_lp", "nonfinite_guard", "time_limit_lp", "mixed_bounds_lp", "transport_8x12_s0"]


def _lp(name):
    if name in lp_cases.REFERENCE_CASES:
        return lp_cases.REFERENCE_CASES[name]()
    if name.startswith("random_60x150_s"):
        return lpgen.random_lp(int(name.rsplit("s", 1)[1]), 60, 150)
    if name == "random_250x625_s7":
        return lpgen.random_lp(7, 250, 625)
    if name == "transport_8x12_s0":
        return lpgen.transport_lp(0, 8, 12)
    if name == "powerlaw_120x300_s3":
        return lpgen.powerlaw_lp(3, 120, 300, 2.179, 2, 40)
    raise KeyError(name)


def rel(a, b):
    return abs(a - b) / max(1.0, abs(b))


@pytest.mark.parametrize("precision", [1, 2])
@pytest.mark.parametrize("name", WELL)
def test_ipm_matches_oracle(hip, name, precision):
    lp = _lp(name)
    print(f"[case] ipm {name} m={lp.m} n={lp.n} precision={precision}", flush=True)
    ref = O.ipm_solve(lp)
    r = hip.ipm_solve(lp, precision=precision, solver=hip.HN_NE_DENSE)
    assert r.status == ref.status
    assert abs(r.iterations - ref.iterations) <= 1
    assert rel(r.objective_value, ref.objective) <= 1e-8
    if ref.status == "OPTIMAL":
        assert r.rp_norm < 1e-8 and r.rd_norm < 1e-8 and r.mu < 1e-8
        # the reference's absolute test leaves the relative gap near 1e-8 or
        # above; with the relative-gap bar on top every measure on the
        # caller's LP meets 1e-8 and the optimum stays the oracle's
        g = hip.ipm_solve(lp, precision=precision, solver=hip.HN_NE_DENSE, gap_tol=1e-8)
        assert g.status == "OPTIMAL"
        assert g.rel_primal <= 1e-8 and g.rel_dual <= 1e-8 and g.rel_gap <= 1e-8
        assert rel(g.objective_value, ref.objective) <= 1e-8


@pytest.mark.parametrize("name", HARD)
def test_ipm_hard_cases_status(hip, name):
    lp = _lp(name)
    print(f"[case] ipm hard {name}", flush=True)
    ref = O.ipm_solve(lp)
    r = hip.ipm_solve(lp, solver=hip.HN_NE_DENSE)
    assert np.isfinite(r.objective_value) and np.all(np.isfinite(r.primal))
    if ref.status == "OPTIMAL":
        assert r.status == "OPTIMAL"
        assert rel(r.objective_value, ref.objective) <= 1e-6
    else:
        assert r.status != "OPTIMAL" or rel(r.objective_value, ref.objective) <= 1e-4


def test_ipm_reference_control_paths(hip):
    # test_lp.cpp:287-325 and :604-634: interrupt, iteration limit, time limit
    r = hip.ipm_solve(lp_cases.simple_lp(), interrupt_callback=lambda: True)
    assert r.status == "INTERRUPTED"
    r = hip.ipm_solve(lp_cases.simple_lp(), max_iterations=1)
    ref = O.ipm_solve(lp_cases.simple_lp(), max_iterations=1)
    assert r.status == "ITER_LIMIT" == ref.status and r.primal.size == 2
    assert rel(r.objective_value, ref.objective) <= 1e-10
    r = hip.ipm_solve(lp_cases.time_limit_lp(), time_limit_seconds=1e-9)
    assert np.isfinite(r.objective_value) and np.all(np.isfinite(r.primal))
    # known answers: objective 20 at (0, 10), 50, 40, 34
    r = hip.ipm_solve(lp_cases.simple_lp())
    assert r.status == "OPTIMAL" and abs(r.objective_value - 20.0) < 1e-6
    assert np.allclose(r.primal, [0.0, 10.0], atol=1e-6)
    assert abs(hip.ipm_solve(lp_cases.integration_lp()).objective_value - 50.0) < 1e-2
    assert abs(hip.ipm_solve(lp_cases.auto_mode_lp()).objective_value - 40.0) < 1e-6
    assert abs(hip.ipm_solve(lp_cases.forced_mode_lp()).objective_value - 34.0) < 1e-6


def test_ipm_standard_form_matches_oracle(hip):
    # the device iterates on the same scaled standard form: identical x after 1 step
    lp = lp_cases.mixed_bounds_lp()
    ref = O.ipm_solve(lp, max_iterations=1)
    r = hip.ipm_solve(lp, max_iterations=1)
    assert r.x.size == ref.x.size and r.y.size == ref.lam.size
    assert np.allclose(r.x, ref.x, rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("seed", [1, 2])
def test_ipm_single_lp_full_size(hip, seed):
    # the single-LP workload: m=2,000, n=5,000, 5 nonzeros per column
    gold = json.load(open(os.path.join(GOLDEN, "ipm_single_lp.json")))[str(seed)]
    lp = lpgen.random_lp(seed, 2000, 5000)
    print(f"[case] ipm 2000x5000 seed {seed}", flush=True)
    for precision in (1, 2):
        r = hip.ipm_solve(lp, precision=precision)
        print(f"  precision={precision} {r} fp64_from={r.fp64_from_iter}", flush=True)
        assert r.status == gold["status"] == "OPTIMAL"
        assert rel(r.objective_value, gold["objective"]) <= 1e-8
        # the reference's slow primal tail around 1e-8 makes the last few
        # steps sensitive to rounding; the count agrees to within 10%
        assert abs(r.iterations - gold["iterations"]) <= max(2, gold["iterations"] // 10)
        assert r.rel_primal <= 1e-8 and r.rel_dual <= 1e-8 and r.rel_gap <= 1e-8


def test_ipm_pcg_path(hip):
    lp = lpgen.random_lp(3, 150, 375)
    ref = O.ipm_solve(lp)
    r = hip.ipm_solve(lp, solver=hip.HN_NE_PCG, pcg_max_iter=20000, gap_tol=1e-8)
    print(f"  pcg {r} cg={r.cg_iterations} rp={r.rp_norm:.2e} rel {r.rel_primal:.1e} {r.rel_dual:.1e} "
          f"{r.rel_gap:.1e}", flush=True)
    assert r.cg_iterations > 0
    assert r.status == ref.status == "OPTIMAL"
    assert rel(r.objective_value, ref.objective) <= 1e-8
    assert r.rel_primal <= 1e-8 and r.rel_dual <= 1e-8 and r.rel_gap <= 1e-8


# ---- HN_PIVOT_SKIP (not in the reference): r = reg, tiny pivots skipped ----
SKIP_CASES = ["random_60x150_s1", "random_250x625_s7", "powerlaw_120x300_s3", "gpu_parity_lp", "transport_8x12_s0",
              "batch_3", "transport_60x90_s1"]


def _lp_skip(name):
    if name == "batch_3":
        return lpgen.batch_lp(3)
    if name == "transport_60x90_s1":
        return lpgen.transport_lp(1, 60, 90)
    return _lp(name)


@pytest.mark.parametrize("precision", [1, 2])
@pytest.mark.parametrize("name", SKIP_CASES)
def test_ipm_skip_rule_matches_oracle(hip, name, precision):
    lp = _lp_skip(name)
    print(f"[case] ipm skip {name} m={lp.m} n={lp.n} precision={precision}", flush=True)
    ref = O.ipm_solve(lp, pivot_rule=1)
    r = hip.ipm_solve(lp, precision=precision, solver=hip.HN_NE_DENSE, pivot_rule=hip.HN_PIVOT_SKIP)
    print(f"  gpu {r} oracle it={ref.iterations} obj={ref.objective!r}", flush=True)
    assert r.status == ref.status == "OPTIMAL"
    if max(ref.iterations, r.iterations) >= 80:
        # the 8 x 12 transportation LP stalls under this rule too (the oracle
        # exits through the relaxed post-loop test); in that regime the step
        # count is rounding-sensitive, the optimum is not
        assert rel(r.objective_value, ref.objective) <= 1e-6
        return
    assert r.iterations < 100
    if not name.startswith("transport"):
        # the last steps on degenerate (transportation) LPs take tiny dual
        # steps whose count is rounding-sensitive; elsewhere it agrees to 20%
        assert abs(r.iterations - ref.iterations) <= max(3, ref.iterations // 5)
    assert rel(r.objective_value, ref.objective) <= 1e-8
    assert r.rp_norm < 1e-8 and r.rd_norm < 1e-8 and r.mu < 1e-8


@pytest.mark.parametrize("pivot_rule", [0, 1])
def test_ipm_runs_are_deterministic(hip, pivot_rule):
    # no atomics take part in any value: two solves give the same bits
    for lp in (lpgen.transport_lp(1, 60, 90), lpgen.batch_lp(3)):
        a = hip.ipm_solve(lp, pivot_rule=pivot_rule)
        b = hip.ipm_solve(lp, pivot_rule=pivot_rule)
        assert a.iterations == b.iterations
        assert a.objective_value == b.objective_value
        assert np.array_equal(a.x, b.x)


def test_ipm_skip_rule_full_size_transport(hip):
    # the 200 x 300 transportation LP: the reference rule stalls until the
    # iteration limit, the skip rule meets the 1e-8 test
    lp = lpgen.transport_lp(0)
    print("[case] ipm skip transport 200x300 seed 0", flush=True)
    ref = O.ipm_solve(lp, pivot_rule=1)
    r = hip.ipm_solve(lp, pivot_rule=hip.HN_PIVOT_SKIP)
    print(f"  gpu {r} oracle it={ref.iterations}", flush=True)
    assert r.status == "OPTIMAL" and r.iterations < 40
    assert abs(r.iterations - ref.iterations) <= max(2, ref.iterations // 10)
    assert r.rp_norm < 1e-8 and r.rd_norm < 1e-8 and r.mu < 1e-8
    assert rel(r.objective_value, ref.objective) <= 1e-8


@pytest.mark.parametrize("name", ["random_60x150_s2", "powerlaw_120x300_s3"])
def test_ipm_pcg_mixed_precision(hip, name):
    # FP32 Jacobi-PCG bulk with FP64 refinement of every Newton step; FP64
    # PCG once mu < mu_switch
    lp = _lp(name)
    ref = O.ipm_solve(lp)
    r = hip.ipm_solve(lp, solver=hip.HN_NE_PCG, precision=hip.HN_PREC_MIXED, pcg_max_iter=20000, gap_tol=1e-8)
    print(f"  mixed pcg {r} cg={r.cg_iterations} sweeps={r.refine_sweeps} fp64_from={r.fp64_from_iter} "
          f"rel {r.rel_primal:.1e} {r.rel_dual:.1e} {r.rel_gap:.1e}", flush=True)
    assert r.status == ref.status == "OPTIMAL"
    assert r.cg_iterations > 0 and r.fp64_from_iter > 0
    assert rel(r.objective_value, ref.objective) <= 1e-8
    assert r.rel_primal <= 1e-8 and r.rel_dual <= 1e-8 and r.rel_gap <= 1e-8


@pytest.mark.parametrize("pivot_rule", [0, 1])
def test_ipm_gap_tol_full_size(hip, pivot_rule):
    # the 2,000 x 5,000 random LP (seed 1) with the relative-gap bar on top of
    # the reference's absolute stop
    lp = lpgen.random_lp(1, 2000, 5000)
    print(f"[case] gap_tol 2000x5000 pivot_rule={pivot_rule}", flush=True)
    r = hip.ipm_solve(lp, pivot_rule=pivot_rule, gap_tol=1e-8)
    p = hip.ipm_solve(lp, pivot_rule=pivot_rule)
    assert r.status == p.status == "OPTIMAL"
    assert r.rel_gap <= 1e-8 and r.rel_primal <= 1e-8 and r.rel_dual <= 1e-8
    assert r.iterations >= p.iterations
    assert abs(r.objective_value - p.objective_value) <= 1e-7 * max(1.0, abs(p.objective_value))
