"""Integer and order-free pieces of the IPM on the device, bit for bit:
the pattern of A D A' (build_normal_equations, interior_point.cpp:144-200)
and the ratio test (max_step, :134-142).  Also the device solver's
per-iteration trace against the reference's own printed trace
(HYPERNOVA_IPM_DBG, :710-714) kept in tests/golden/ipm_reference.json."""
import json
import os
import re

import numpy as np
import pytest

import lp_cases
import lpgen
import oracle as O

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _std(hip, lp):
    return hip.lp_standard_form(lp)


PATTERN_CASES = ["random_60x150_s1", "powerlaw_120x300_s3", "mixed_bounds_lp", "transport_8x12_s0",
                 "random_2000x5000_s1", "powerlaw_2000x5000_s3", "empty_rows"]


def _pattern_input(hip, name):
    if name == "empty_rows":
        # rows 0, 5, 6 and the last are empty; columns of 1..40 entries
        rng = np.random.default_rng(3)
        m, n = 300, 120
        cols = [np.sort(rng.choice(np.setdiff1d(np.arange(m), [0, 5, 6, m - 1]), size=int(k), replace=False))
                for k in rng.integers(1, 41, n)]
        rows = [[] for _ in range(m)]
        for j, rs in enumerate(cols):
            for i in rs:
                rows[i].append(j)
        rp = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int64)
        ci = np.concatenate([np.array(r, np.int64) for r in rows])
        return m, n, rp, ci, rng.standard_normal(ci.size)
    if name in lp_cases.REFERENCE_CASES:
        lp = lp_cases.REFERENCE_CASES[name]()
    elif name.startswith("random_"):
        dims, seed = name.split("_")[1], int(name.rsplit("s", 1)[1])
        m, n = (int(t) for t in dims.split("x"))
        lp = lpgen.random_lp(seed, m, n)
    elif name.startswith("powerlaw_"):
        dims = name.split("_")[1]
        m, n = (int(t) for t in dims.split("x"))
        lp = lpgen.powerlaw_lp(3, m, n, 2.179, 2, 40 if m < 1000 else 200)
    else:
        lp = lpgen.transport_lp(0, 8, 12)
    f = _std(hip, lp)
    return f["m"], f["n"], f["rp"], f["ci"], f["v"]


@pytest.mark.parametrize("name", PATTERN_CASES)
def test_adat_pattern_bit_exact(hip, name):
    m, n, rp, ci, v = _pattern_input(hip, name)
    print(f"[case] A D A' pattern {name} m={m} n={n} nnz={ci.size}", flush=True)
    A = hip.CsrMatrix.upload(v, ci, rp, n)
    N = A.adat_pattern()
    nv, nc, nr = N.download()
    orp, oci = O.adat_pattern(rp, ci, m, n)
    assert N.nrows == m
    assert np.array_equal(nr.astype(np.int64), orp)
    assert np.array_equal(nc.astype(np.int64), oci)
    assert not nv.any()


@pytest.mark.parametrize("n", [0, 1, 63, 1000, 100_003])
def test_ratio_test_bit_exact(hip, n):
    rng = np.random.default_rng(n + 1)
    x, z = rng.uniform(1e-9, 3.0, n), rng.uniform(1e-9, 3.0, n)
    dx, dz = rng.standard_normal(n) * 10.0 ** rng.uniform(-3, 3, n), rng.standard_normal(n)
    if n > 10:
        dz[:] = np.abs(dz)  # no dz component blocks: the step is 0.99
        dx[n // 3] = -x[n // 3] * 1e6  # one tiny blocking step
    out = hip.DeviceArray(2)
    d = [hip.DeviceArray.from_numpy(a) for a in (x, dx, z, dz)]
    hip.check(hip.LIB.hn_ratio_test_dev(d[0].ptr, d[1].ptr, d[2].ptr, d[3].ptr, n, out.ptr, None))
    got = out.to_numpy()
    want = np.array([O.max_step(x, dx), O.max_step(z, dz)])
    print(f"[case] ratio test n={n}: {got!r} vs {want!r}", flush=True)
    assert got.tobytes() == want.tobytes()


TRACE_CASES = ["simple_lp", "integration_lp", "gpu_parity_lp", "auto_mode_lp", "forced_mode_lp",
               "random_60x150_s1", "random_60x150_s2", "random_60x150_s3", "random_60x150_s4", "random_60x150_s5",
               "random_250x625_s7", "powerlaw_120x300_s3"]
LINE = re.compile(r"IPM it=(\d+) mu=(\S+) rp=(\S+) rd=(\S+) ap=(\S+) ad=(\S+)")


def _trace_lp(name):
    if name in lp_cases.REFERENCE_CASES:
        return lp_cases.REFERENCE_CASES[name]()
    if name.startswith("random_60x150_s"):
        return lpgen.random_lp(int(name.rsplit("s", 1)[1]), 60, 150)
    if name == "random_250x625_s7":
        return lpgen.random_lp(7, 250, 625)
    return lpgen.powerlaw_lp(3, 120, 300, 2.179, 2, 40)


@pytest.mark.parametrize("name", TRACE_CASES)
def test_trace_matches_reference(hip, capfd, name):
    """The reference prints mu, ||rp||, ||rd||, ap, ad with %g (6 digits) per
    iteration.  Its factor is a minimum-degree sparse LDL', the device's a
    dense natural-order LDL' of the same matrix, so the steps agree to
    rounding: every printed value must agree to 1e-4 relative, or to 1e-10
    of its first printed value (a residual near zero is a difference of
    rounded terms), while mu is above 1e-6 (later iterates sit at the
    ill-conditioned end, where the last digits of the two factorisations
    part), and the iteration count exactly."""
    gold = json.load(open(os.path.join(GOLDEN, "ipm_reference.json")))[name]
    ref = [list(map(float, t)) for t in (gold["trace"] if isinstance(gold["trace"], list)
                                         else json.loads(gold["trace"]))]
    r = hip.ipm_solve(_trace_lp(name), solver=hip.HN_NE_DENSE, verbose=True)
    err = capfd.readouterr().err
    got = [[float(v) for v in m.groups()[1:]] for m in LINE.finditer(err)]
    print(f"[case] trace {name}: {len(got)} device lines, {len(ref)} reference lines, it {r.iterations}",
          flush=True)
    assert r.iterations == int(gold["iterations"]) == len(ref)
    assert len(got) == len(ref)
    for k, (g, w) in enumerate(zip(got, ref)):
        if w[0] < 1e-6:
            break
        for a, b, b0 in zip(g, w, ref[0]):
            assert abs(a - b) <= 1e-4 * abs(b) + 1e-10 * abs(b0), (name, k, g, w)
