"""The three ways hn_ipm_solve drives an iteration (hn_ipm_opts.loop) give
the same iterates.

HN_LOOP_HOST decides sigma and the refinement sweeps on the host;
HN_LOOP_DEVICE keeps an iteration on the device between two reads of its
scalars (sigma, the refinement checks and the failure flags formed there);
HN_LOOP_GRAPH also replays the iteration as a captured HIP graph.  The
kernels and their order are the same, so status, iteration count, refinement
sweeps and the final point agree; sigma's cube is formed in double-double on
the device and by std::pow on the host, which round alike except in near-tie
cases, so the point is compared to 1e-12 rather than bit for bit.  An
iteration whose solve needs more refinement sweeps than the device step
holds is undone and run on the host path (host_redos), which keeps the
iterates identical too: stair and sctap1 under the skip rule need sweeps,
and exercise that path."""
import json
import os

import numpy as np
import pytest

import lpgen
import mps

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
FIX = json.load(open(os.path.join(HERE, "golden", "netlib.json")))


def _netlib(name):
    return mps.read_mps(os.path.join(HERE, "golden", "netlib", name + ".mps.gz"))[0]


CASES = {
    "afiro": lambda: (_netlib("afiro"), {}),
    "blend": lambda: (_netlib("blend"), {}),
    "sctap1": lambda: (_netlib("sctap1"), {}),
    "sctap1_skip": lambda: (_netlib("sctap1"), {"pivot_rule": 1}),
    "bandm": lambda: (_netlib("bandm"), {}),
    "stair": lambda: (_netlib("stair"), {}),
    "25fv47": lambda: (_netlib("25fv47"), {}),
    "random_300x700": lambda: (lpgen.random_lp(4, 300, 700), {}),
    "transport_20x30": lambda: (lpgen.transport_lp(5, 20, 30), {}),
    "transport_20x30_skip": lambda: (lpgen.transport_lp(5, 20, 30), {"pivot_rule": 1}),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_loop_modes_agree(hip, name):
    lp, kw = CASES[name]()
    out = {}
    for mode in (hip.HN_LOOP_HOST, hip.HN_LOOP_DEVICE, hip.HN_LOOP_GRAPH):
        r = hip.ipm_solve(lp, loop=mode, **kw)
        print(f"[case] {name} loop={mode}: {r.status} it={r.iterations} sweeps={r.refine_sweeps} "
              f"device_steps={r.device_steps} host_redos={r.host_redos} {r.solve_time_ms:.1f} ms", flush=True)
        out[mode] = r
    h = out[hip.HN_LOOP_HOST]
    assert h.device_steps == 0
    for mode in (hip.HN_LOOP_DEVICE, hip.HN_LOOP_GRAPH):
        r = out[mode]
        assert r.device_steps > 0
        assert r.status == h.status and r.iterations == h.iterations
        assert r.ne_solves == h.ne_solves and r.refine_sweeps == h.refine_sweeps
        scale = max(1.0, np.abs(h.x).max())
        assert np.abs(r.x - h.x).max() <= 1e-12 * scale
        assert abs(r.objective_value - h.objective_value) <= 1e-12 * max(1.0, abs(h.objective_value))
        assert (r.host_redos > 0) == (h.refine_sweeps > 0) and r.host_redos <= 1
    if name in ("afiro", "blend", "sctap1", "25fv47"):  # the reference IPM's own status and count
        d = FIX[name]
        assert out[hip.HN_LOOP_GRAPH].status == d["ref_ipm_status"]
        assert out[hip.HN_LOOP_GRAPH].iterations == d["ref_ipm_iterations"]


def test_loop_graph_with_interrupt_and_iteration_limit(hip):
    # the host's tests between iterations still apply to replayed steps
    lp = _netlib("sctap1")
    r = hip.ipm_solve(lp, max_iterations=5)
    assert r.status == "ITER_LIMIT" and r.iterations == 5 and r.device_steps == 5
    calls = []

    def stop():
        calls.append(1)
        return len(calls) > 7

    r = hip.ipm_solve(lp, interrupt_callback=stop)
    assert r.status == "INTERRUPTED" and r.iterations == 7


def test_loop_rejects_bad_mode(hip):
    with pytest.raises(RuntimeError):
        hip.ipm_solve(lpgen.random_lp(1, 20, 50), loop=7)
