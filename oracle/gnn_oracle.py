"""CPU restatement of the warm-start graph network (csrc/hn_gnn.hip).

TEST INFRASTRUCTURE ONLY: tests/ use it as the checker of the device
forward pass and gradients; the product never imports it.

The reference has no learned component (it only accepts a start point,
core/lp/interior_point.cpp:443-453), so there is nothing to restate from it:
this file states the network the device implements, and the finite-difference
test in tests/test_gnn.py pins its backward pass (exact=True switches off the
BF16 rounding and runs in float64).

Network (H = 64, 4 layers) on an LP's scaled standard form (M: m x n):
  features  variable j: c_j/cbar, [deg_j == 1], 1, min/max over j's rows of
            rank(c_j within the row), min/max of (c_j - row min)/cbar;
            constraint i: b_i/bbar, [row i holds a degree-1 column], 1
  embed     hv0 = relu(Fv Wev + bev), hc0 = relu(Fc Wec + bec)
  layer l   ac = Acv hv_l (Acv_ij = M_ij/deg_i); hc_{l+1} = relu([hc_l, ac] Wc_l + bc_l)
            av = Avc hc_{l+1} (Avc_ji = M_ij/deg_j); hv_{l+1} = relu([hv_l, av] Wv_l + bv_l)
  heads     log x = hv4 wx + bx, log s = hv4 ws + bs, y = (hc4 wy + by) cbar
  loss      mean (log x - tx)^2 + mean (log s - ts)^2 + 1e-3 mean ((y - ty)/cbar)^2
MFMA operands are BF16 (round to nearest even): the GEMM inputs [h, agg] and
W forward; the masked output gradient, W and [h, agg] backward.
"""
from __future__ import annotations

import numpy as np

H, NL, FV, FC = 64, 4, 7, 3
YW = 1e-3


def layout():
    o = 0
    L = {}

    def take(name, n):
        nonlocal o
        L[name] = (o, n)
        o += n

    take("wev", FV * H)
    take("bev", H)
    take("wec", FC * H)
    take("bec", H)
    for l in range(NL):
        take(f"wc{l}", 2 * H * H)
        take(f"bc{l}", H)
        take(f"wv{l}", 2 * H * H)
        take(f"bv{l}", H)
    for h in ("x", "s", "y"):
        take(f"w{h}", H)
        take(f"b{h}", 1)
    L["total"] = (o, 0)
    return L


LAYOUT = layout()
NPARAMS = LAYOUT["total"][0]


def unpack(p):
    out = {}
    for k, (o, n) in LAYOUT.items():
        if k == "total":
            continue
        out[k] = p[o:o + n]
    for k in list(out):
        if k in ("wev",):
            out[k] = out[k].reshape(FV, H)
        elif k == "wec":
            out[k] = out[k].reshape(FC, H)
        elif k.startswith("wc") or k.startswith("wv"):
            out[k] = out[k].reshape(2 * H, H)
    return out


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    x = np.ascontiguousarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    r = np.where(nan, ((u >> 16) | 0x40).astype(np.uint32) << 16, r)
    return r.view(np.float32)


class Graph:
    def __init__(self, m, n, rp, ci):
        self.m, self.n = int(m), int(n)
        self.rp = np.asarray(rp, np.int64)
        self.ci = np.asarray(ci, np.int64)
        self.rows = np.repeat(np.arange(self.m), np.diff(self.rp))
        self.deg_r = np.diff(self.rp)
        self.deg_c = np.bincount(self.ci, minlength=self.n)


def features(g: Graph, val, b, c, dt=np.float32):
    """The device's feature pipeline (k_gnn_scales .. k_gnn_feat_c)."""
    c = np.asarray(c, np.float64)
    b = np.asarray(b, np.float64)
    mc = np.abs(c).mean() if g.n else 0.0
    mb = np.abs(b).mean() if g.m else 0.0
    cbar = np.float32(mc if mc > 0 else 1.0)
    bbar = np.float32(mb if mb > 0 else 1.0)
    nnz = g.ci.size
    rank = np.zeros(nnz, np.float32)
    gap = np.zeros(nnz, np.float32)
    for i in range(g.m):
        k0, k1 = g.rp[i], g.rp[i + 1]
        if k1 == k0:
            continue
        js = g.ci[k0:k1]
        cs = c[js]
        mn = cs.min()
        gap[k0:k1] = ((cs - mn) / np.float64(cbar)).astype(np.float32)
        if k1 - k0 > 4096:
            rank[k0:k1] = 0.5
            continue
        order = np.lexsort((js, cs))          # by cost, ties by column index
        r = np.empty(k1 - k0, np.int64)
        r[order] = np.arange(k1 - k0)
        rank[k0:k1] = r.astype(np.float32) / np.float32(k1 - k0)
    Fv = np.zeros((g.n, FV), dt)
    Fv[:, 0] = (c / np.float64(cbar)).astype(np.float32)
    Fv[:, 1] = (g.deg_c == 1)
    Fv[:, 2] = 1.0
    big = np.float32(np.inf)
    rmin = np.full(g.n, big, np.float32)
    rmax = np.full(g.n, -big, np.float32)
    gmin = np.full(g.n, big, np.float32)
    gmax = np.full(g.n, -big, np.float32)
    np.minimum.at(rmin, g.ci, rank)
    np.maximum.at(rmax, g.ci, rank)
    np.minimum.at(gmin, g.ci, gap)
    np.maximum.at(gmax, g.ci, gap)
    empty = g.deg_c == 0
    for a in (rmin, rmax, gmin, gmax):
        a[empty] = 0.0
    Fv[:, 3], Fv[:, 4], Fv[:, 5], Fv[:, 6] = rmin, rmax, gmin, gmax
    Fc = np.zeros((g.m, FC), dt)
    Fc[:, 0] = (b / np.float64(bbar)).astype(np.float32)
    slack_rows = np.zeros(g.m, bool)
    slack_rows[g.rows[g.deg_c[g.ci] == 1]] = True
    Fc[:, 1] = slack_rows
    Fc[:, 2] = 1.0
    return Fv, Fc, cbar, bbar


def _agg_cv(g, val, hv, dt):
    w = (np.asarray(val, np.float64).astype(dt) / np.maximum(g.deg_r[g.rows], 1).astype(dt))
    out = np.zeros((g.m, hv.shape[1]), dt)
    np.add.at(out, g.rows, w[:, None] * hv[g.ci])
    return out


def _agg_vc(g, val, hc, dt):
    w = (np.asarray(val, np.float64).astype(dt) / np.maximum(g.deg_c[g.ci], 1).astype(dt))
    out = np.zeros((g.n, hc.shape[1]), dt)
    np.add.at(out, g.ci, w[:, None] * hc[g.rows])
    return out


def _aggT_cv(g, val, dac, dt):
    """Acv' dac: to variables, weight M_ij / deg_i."""
    w = (np.asarray(val, np.float64).astype(dt) / np.maximum(g.deg_r[g.rows], 1).astype(dt))
    out = np.zeros((g.n, dac.shape[1]), dt)
    np.add.at(out, g.ci, w[:, None] * dac[g.rows])
    return out


def _aggT_vc(g, val, dav, dt):
    """Avc' dav: to constraints, weight M_ij / deg_j."""
    w = (np.asarray(val, np.float64).astype(dt) / np.maximum(g.deg_c[g.ci], 1).astype(dt))
    out = np.zeros((g.m, dav.shape[1]), dt)
    np.add.at(out, g.rows, w[:, None] * dav[g.ci])
    return out


def forward(params, g: Graph, val, b, c, exact=False):
    """Activations of the forward pass (dict).  exact=True: float64, no BF16."""
    dt = np.float64 if exact else np.float32
    rnd = (lambda a: a) if exact else bf16
    P = {k: np.asarray(v, dt) for k, v in unpack(np.asarray(params, np.float64 if exact else np.float32)).items()}
    Fv, Fc, cbar, bbar = features(g, val, b, c, dt)
    act = dict(Fv=Fv, Fc=Fc, cbar=dt(cbar))
    hv = np.maximum(Fv @ P["wev"] + P["bev"], 0)
    hc = np.maximum(Fc @ P["wec"] + P["bec"], 0)
    act["hv0"], act["hc0"] = hv, hc
    for l in range(NL):
        ac = _agg_cv(g, val, hv, dt)
        Xc = np.concatenate([hc, ac], 1)
        hc = np.maximum((rnd(Xc).astype(np.float64) @ rnd(P[f"wc{l}"]).astype(np.float64)).astype(dt) + P[f"bc{l}"], 0)
        av = _agg_vc(g, val, hc, dt)
        Xv = np.concatenate([hv, av], 1)
        hv = np.maximum((rnd(Xv).astype(np.float64) @ rnd(P[f"wv{l}"]).astype(np.float64)).astype(dt) + P[f"bv{l}"], 0)
        act[f"ac{l}"], act[f"av{l}"] = ac, av
        act[f"hc{l + 1}"], act[f"hv{l + 1}"] = hc, hv
    act["lx"] = hv @ P["wx"] + P["bx"][0]
    act["ls"] = hv @ P["ws"] + P["bs"][0]
    act["ly"] = (hc @ P["wy"] + P["by"][0]) * act["cbar"]
    return act


def loss_and_grad(params, g: Graph, val, b, c, tx, ts, ty, exact=False):
    """Loss and the flat gradient, by the same backward pass as the device."""
    dt = np.float64 if exact else np.float32
    rnd = (lambda a: a) if exact else bf16
    P = {k: np.asarray(v, dt) for k, v in unpack(np.asarray(params, np.float64 if exact else np.float32)).items()}
    A = forward(params, g, val, b, c, exact)
    n, m = g.n, g.m
    cbar = A["cbar"]
    ex = A["lx"] - tx
    es = A["ls"] - ts
    ey = (A["ly"] - ty) / cbar
    loss = (ex * ex).mean() + (es * es).mean() + YW * (ey * ey).mean()
    G = {}
    dlx = 2 * ex / n
    dls = 2 * es / n
    dly = 2 * YW * ey / m             # gradient of the pre-scaling head output
    hv, hc = A[f"hv{NL}"], A[f"hc{NL}"]
    G["wx"], G["bx"] = hv.T @ dlx, np.array([dlx.sum()])
    G["ws"], G["bs"] = hv.T @ dls, np.array([dls.sum()])
    G["wy"], G["by"] = hc.T @ dly, np.array([dly.sum()])
    dhv = np.outer(dlx, P["wx"]) + np.outer(dls, P["ws"])
    dhc = np.outer(dly, P["wy"])
    for l in range(NL - 1, -1, -1):
        # variable update
        Y = A[f"hv{l + 1}"]
        dZ = np.where(Y > 0, dhv, 0).astype(dt)
        X = np.concatenate([A[f"hv{l}"], A[f"av{l}"]], 1)
        G[f"wv{l}"] = (rnd(X).astype(np.float64).T @ rnd(dZ).astype(np.float64)).astype(dt)
        G[f"bv{l}"] = dZ.sum(0)
        dX = (rnd(dZ).astype(np.float64) @ rnd(P[f"wv{l}"]).astype(np.float64).T).astype(dt)
        dhv_prev, dav = dX[:, :H], dX[:, H:]
        dhc = dhc + _aggT_vc(g, val, dav, dt)
        # constraint update
        Y = A[f"hc{l + 1}"]
        dZ = np.where(Y > 0, dhc, 0).astype(dt)
        X = np.concatenate([A[f"hc{l}"], A[f"ac{l}"]], 1)
        G[f"wc{l}"] = (rnd(X).astype(np.float64).T @ rnd(dZ).astype(np.float64)).astype(dt)
        G[f"bc{l}"] = dZ.sum(0)
        dX = (rnd(dZ).astype(np.float64) @ rnd(P[f"wc{l}"]).astype(np.float64).T).astype(dt)
        dhc, dac = dX[:, :H], dX[:, H:]
        dhv = dhv_prev + _aggT_cv(g, val, dac, dt)
    dZ = np.where(A["hv0"] > 0, dhv, 0)
    G["wev"], G["bev"] = A["Fv"].T @ dZ, dZ.sum(0)
    dZ = np.where(A["hc0"] > 0, dhc, 0)
    G["wec"], G["bec"] = A["Fc"].T @ dZ, dZ.sum(0)
    flat = np.zeros(NPARAMS, np.float64)
    for k, (o, nn) in LAYOUT.items():
        if k == "total":
            continue
        flat[o:o + nn] = np.asarray(G[k], np.float64).ravel()
    return float(loss), flat


def init_params(seed: int) -> np.ndarray:
    """The device's Glorot-uniform initialisation (splitmix64 stream)."""
    mask = (1 << 64) - 1
    state = [seed & mask]

    def nxt():
        state[0] = (state[0] + 0x9E3779B97F4A7C15) & mask
        z = state[0]
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        return z ^ (z >> 31)

    p = np.zeros(NPARAMS, np.float32)

    def fill(name, fin, fout):
        o, n = LAYOUT[name]
        a = np.sqrt(6.0 / (fin + fout))
        for e in range(fin * fout):
            u = (nxt() >> 11) * (1.0 / 9007199254740992.0)
            p[o + e] = np.float32((2.0 * u - 1.0) * a)

    fill("wev", FV, H)
    fill("wec", FC, H)
    for l in range(NL):
        fill(f"wc{l}", 2 * H, H)
        fill(f"wv{l}", 2 * H, H)
    fill("wx", H, 1)
    fill("ws", H, 1)
    fill("wy", H, 1)
    return p
