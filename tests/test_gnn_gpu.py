"""Warm-start network kernels (hn_gnn_*) against the CPU oracle
(oracle/gnn_oracle.py, which emulates the same BF16 rounding points).

Tolerances: the oracle sums in a different order than the device, so a
value feeding a BF16 rounding can land on the other side of a rounding
boundary; results agree to about 1e-2 relative in norm, not bit for bit."""
import numpy as np
import pytest

import gnn_oracle as G
import hypernova_hip as HH
import lpgen

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ws(hip):
    import warmstart
    return warmstart


def _case(ws, seed=0, S=12, T=18):
    lp = lpgen.transport_lp(seed, S, T)
    sf = HH.lp_standard_form(lp)
    g = ws.Graph(sf["m"], sf["n"], sf["rp"], sf["ci"])
    og = G.Graph(sf["m"], sf["n"], sf["rp"], sf["ci"])
    dev = {k: HH.DeviceArray.from_numpy(sf[k]) for k in ("v", "rhs", "c")}
    return lp, sf, g, og, dev


def relnorm(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def test_gnn_init_matches_oracle(ws):
    net = ws.Network(seed=11)
    assert net.nparams == G.NPARAMS
    assert np.array_equal(net.get_params(), G.init_params(11))


def test_gnn_forward_matches_oracle(ws):
    net = ws.Network(seed=3)
    lp, sf, g, og, dev = _case(ws)
    print(f"[case] gnn forward transport 12x18 m={sf['m']} n={sf['n']}", flush=True)
    lx, ls, ly = net.forward(g, dev["v"], dev["rhs"], dev["c"])
    act = G.forward(net.get_params(), og, sf["v"], sf["rhs"], sf["c"])
    assert relnorm(lx, act["lx"]) < 1e-2 and relnorm(ls, act["ls"]) < 1e-2 and relnorm(ly, act["ly"]) < 1e-2


def test_gnn_gradients_match_oracle(ws):
    net = ws.Network(seed=5)
    lp, sf, g, og, dev = _case(ws, seed=1)
    rng = np.random.default_rng(0)
    tx = rng.normal(size=sf["n"]).astype(np.float32)
    ts = rng.normal(size=sf["n"]).astype(np.float32)
    ty = rng.normal(size=sf["m"]).astype(np.float32)
    print(f"[case] gnn gradients transport 12x18", flush=True)
    loss, grad = net.gradients(g, dev["v"], dev["rhs"], dev["c"], HH.DeviceArray.from_numpy(tx),
                               HH.DeviceArray.from_numpy(ts), HH.DeviceArray.from_numpy(ty))
    ref_loss, ref = G.loss_and_grad(net.get_params(), og, sf["v"], sf["rhs"], sf["c"], tx, ts, ty)
    assert abs(loss - ref_loss) <= 1e-3 * abs(ref_loss)
    for name, (o, nn) in G.LAYOUT.items():
        if name == "total":
            continue
        assert relnorm(grad[o:o + nn], ref[o:o + nn]) < 3e-2, name


def test_gnn_training_lowers_loss(ws):
    # eight instances of a small family, fitted to their own IPM targets
    def make(s):
        return lpgen.transport_lp(s, 12, 18)
    ds = ws.make_dataset(make, range(8), chunk=8, log=lambda *a: None)
    net = ws.Network(seed=1)
    hist = ws.train(net, ds, epochs=30, lr=3e-3, lr_final=1e-3, log=lambda *a: None)
    print(f"  losses {hist[0]:.4f} -> {hist[-1]:.4f}", flush=True)
    assert np.all(np.isfinite(hist)) and hist[-1] < 0.5 * hist[0]


def test_trained_warm_start_saves_iterations(ws):
    # the network trained on 20,000 plain-IPM targets of the 200 x 300
    # transportation family (data/warmstart_transport_200x300.npy), applied
    # to held-out seeds: same optimum, fewer iterations than the standard start
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(ws.__file__)), "data", "warmstart_transport_200x300.npy")
    net = ws.Network(0)
    net.set_params(np.load(path))
    print("[case] trained warm start, 64 held-out 200x300 transportation LPs", flush=True)
    r = ws.evaluate(net, ws.transport_family(), range(1_000_000, 1_000_064), log=lambda *a: None)
    print(f"  {r}", flush=True)
    assert r["cold_optimal"] == r["warm_optimal"] == 64
    assert r["accepted"] >= 60
    assert r["max_rel_objective_diff"] <= 1e-6
    assert r["warm_iterations_median"] <= r["cold_iterations_median"] - 3


def test_device_warm_start_in_batch_matches_host_prediction(ws):
    # hn_ipm_solve_batch_gnn predicts each start on the device from the
    # standard form the solver packed; it must give the same solves as the
    # host path (predict(), then the points passed in).  An LP of another
    # size in the same batch has another pattern: it starts cold and solves
    # exactly as a cold solve does.
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(ws.__file__)), "data", "warmstart_transport_200x300.npy")
    net = ws.Network(0)
    net.set_params(np.load(path))
    lps = [lpgen.transport_lp(s, 200, 300) for s in range(1_000_100, 1_000_124)]
    odd = lpgen.transport_lp(5, 20, 30)
    sf = HH.lp_standard_form(lps[0])
    g = ws.Graph(sf["m"], sf["n"], sf["rp"], sf["ci"])
    print("[case] device warm start in the batch, 24 + 1 LPs", flush=True)
    dev = HH.ipm_solve_batch(lps + [odd], pivot_rule=HH.HN_PIVOT_SKIP, gnn=(net.h, g.h))
    host = HH.ipm_solve_batch(lps, pivot_rule=HH.HN_PIVOT_SKIP, warm=ws.predict(net, lps, g))
    cold_odd = HH.ipm_solve_batch([odd], pivot_rule=HH.HN_PIVOT_SKIP)[0]
    accepted = 0
    for k, (d, h) in enumerate(zip(dev, host)):
        assert d.status == h.status == "OPTIMAL", k
        assert d.warm_started == h.warm_started, (k, d.warm_started, h.warm_started)
        accepted += d.warm_started == 1
        assert abs(d.iterations - h.iterations) <= 1, (k, d.iterations, h.iterations)
        assert abs(d.objective_value - h.objective_value) <= 1e-7 * max(1.0, abs(h.objective_value)), k
    assert accepted >= 20
    d = dev[-1]
    assert d.warm_started == 0 and d.status == cold_odd.status == "OPTIMAL"
    assert d.iterations == cold_odd.iterations and d.objective_value == cold_odd.objective_value
