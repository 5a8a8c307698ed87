"""Pins the warm-start network's CPU oracle (oracle/gnn_oracle.py) before it
is trusted as the checker of the device kernels: BF16 rounding on known
values, and the backward pass against central finite differences of the
float64 loss (exact=True: no rounding), on small random graphs."""
import numpy as np
import pytest

import gnn_oracle as G


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 0.0, 65504.0], np.float32)
    r = G.bf16(x)
    # 1+2^-8 is a tie between 1 and 1+2^-7: nearest even is 1; 1+3*2^-8 rounds up to 1+2^-6
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == np.float32(1.0 + 2 ** -6) and r[3] == 1.0
    assert r[4] == -2.5 and r[5] == 0.0 and r[6] == 65536.0
    assert np.isnan(G.bf16(np.array([np.nan], np.float32)))[0]


def _graph(seed, m=7, n=13, k=3):
    rng = np.random.default_rng(seed)
    rows = []
    for j in range(n):
        rows.append(rng.choice(m, size=min(k, m), replace=False))
    rp = [0]
    ci = []
    for i in range(m):
        cols = sorted(j for j in range(n) if i in rows[j])
        ci += cols
        rp.append(len(ci))
    g = G.Graph(m, n, np.array(rp), np.array(ci))
    val = rng.uniform(0.5, 2.0, len(ci))
    b = rng.uniform(1, 5, m)
    c = rng.uniform(0.1, 3, n)
    return g, val, b, c, rng


def test_param_layout():
    assert G.NPARAMS == G.FV * G.H + G.H + G.FC * G.H + G.H + G.NL * 2 * (2 * G.H * G.H + G.H) + 3 * (G.H + 1)


@pytest.mark.parametrize("seed", [1, 2])
def test_oracle_gradient_against_finite_differences(seed):
    g, val, b, c, rng = _graph(seed)
    p = G.init_params(seed).astype(np.float64)
    # biases away from zero so that few units sit on a ReLU kink
    for name, (o, nn) in G.LAYOUT.items():
        if name.startswith("b") and name != "total":
            p[o:o + nn] = rng.uniform(0.05, 0.3, nn)
    tx = rng.normal(size=g.n)
    ts = rng.normal(size=g.n)
    ty = rng.normal(size=g.m)
    loss, grad = G.loss_and_grad(p, g, val, b, c, tx, ts, ty, exact=True)
    assert np.isfinite(loss) and np.abs(grad).max() > 0
    idx = np.concatenate([rng.choice(G.NPARAMS, 120, replace=False),
                          [G.LAYOUT["bx"][0], G.LAYOUT["by"][0], G.LAYOUT["wev"][0], G.LAYOUT["wc3"][0] + 5]])
    eps = 1e-6
    bad = 0
    for i in idx:
        q = p.copy()
        q[i] += eps
        lp, _ = G.loss_and_grad(q, g, val, b, c, tx, ts, ty, exact=True)
        q[i] -= 2 * eps
        lm, _ = G.loss_and_grad(q, g, val, b, c, tx, ts, ty, exact=True)
        fd = (lp - lm) / (2 * eps)
        if abs(fd - grad[i]) > 1e-5 * max(1.0, abs(grad[i])):
            bad += 1
    # a coordinate can straddle a ReLU kink; nearly all must agree
    assert bad <= 2, bad


def test_oracle_features_closed_form():
    g, val, b, c, _ = _graph(3)
    Fv, Fc, cbar, bbar = G.features(g, val, b, c)
    assert np.isclose(cbar, np.abs(c).mean()) and np.isclose(bbar, np.abs(b).mean())
    for j in range(g.n):
        rows = [i for i in range(g.m) if j in g.ci[g.rp[i]:g.rp[i + 1]]]
        ranks = []
        for i in rows:
            cols = g.ci[g.rp[i]:g.rp[i + 1]]
            ranks.append(sum(1 for jj in cols if (c[jj], jj) < (c[j], j)) / len(cols))
        assert np.isclose(Fv[j, 3], min(ranks)) and np.isclose(Fv[j, 4], max(ranks))
