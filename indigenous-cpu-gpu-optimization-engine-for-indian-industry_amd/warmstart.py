"""Learned warm start for the device IPM (host side of csrc/hn_gnn.hip).

A bipartite constraint-variable graph network (4 layers, hidden 64, BF16 MFMA)
predicts (log x, log s, y) on an LP's scaled standard form.  It is trained on
points the plain IPM itself produces: the first iterate of a cold solve whose
mu is at most ``MU_TARGET`` (the solver's ``mu_target`` stop).  That target is
deliberate - a warm start at the predicted *optimum* sits on the boundary of
the positive orthant, and the reference's Mehrotra steps from there stall
under small prediction errors, while a point on the central path at a
moderate mu tolerates them (an error of 20% in x and s still saves most of
the iterations it skips).

The IPM accepts the prediction only if it passes the solver's centrality
test (hn_ipm_solve_batch); otherwise it starts from the standard point.

Nothing here runs on the CPU: every computation of the network is a device
kernel behind libhn_kernels.so.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

import hypernova_hip as HH

MU_TARGET = 5e-2
_dp, _fp, _vp = C.POINTER(C.c_double), C.POINTER(C.c_float), C.c_void_p
_i64p = C.POINTER(C.c_int64)

_L = HH.LIB
_L.hn_gnn_create.argtypes = [C.c_uint64, C.POINTER(_vp)]
_L.hn_gnn_free.argtypes = [_vp]
_L.hn_gnn_num_params.argtypes = [_vp, C.POINTER(C.c_uint64)]
_L.hn_gnn_get_params.argtypes = [_vp, _fp]
_L.hn_gnn_set_params.argtypes = [_vp, _fp]
_L.hn_gnn_graph_create.argtypes = [C.c_int64, C.c_int64, _i64p, _i64p, C.POINTER(_vp)]
_L.hn_gnn_graph_free.argtypes = [_vp]
_L.hn_gnn_forward_dev.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
_L.hn_gnn_gradients_dev.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _fp, _fp]
_L.hn_gnn_train_step_dev.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, C.c_float, _vp, _vp]
_L.hn_gnn_warm_start_dev.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
for _f in ("hn_gnn_create", "hn_gnn_free", "hn_gnn_num_params", "hn_gnn_get_params", "hn_gnn_set_params",
           "hn_gnn_graph_create", "hn_gnn_graph_free", "hn_gnn_forward_dev", "hn_gnn_gradients_dev",
           "hn_gnn_train_step_dev", "hn_gnn_warm_start_dev"):
    getattr(_L, _f).restype = C.c_int


class Graph:
    """Device copy of a standard-form sparsity pattern (hn_gnn_graph)."""

    def __init__(self, m: int, n: int, row_ptr, col):
        rp = np.ascontiguousarray(row_ptr, np.int64)
        ci = np.ascontiguousarray(col, np.int64)
        h = _vp()
        HH.check(_L.hn_gnn_graph_create(m, n, rp.ctypes.data_as(_i64p), ci.ctypes.data_as(_i64p), C.byref(h)))
        self.h, self.m, self.n, self.nnz = h, int(m), int(n), int(ci.size)
        self.row_ptr, self.col = rp, ci

    def __del__(self):
        try:
            if self.h:
                _L.hn_gnn_graph_free(self.h)
        except Exception:
            pass


class Network:
    """Parameters and Adam state of the warm-start network (hn_gnn)."""

    def __init__(self, seed: int = 0):
        h = _vp()
        HH.check(_L.hn_gnn_create(seed, C.byref(h)))
        self.h = h
        n = C.c_uint64()
        HH.check(_L.hn_gnn_num_params(h, C.byref(n)))
        self.nparams = int(n.value)

    def __del__(self):
        try:
            if self.h:
                _L.hn_gnn_free(self.h)
        except Exception:
            pass

    def get_params(self) -> np.ndarray:
        p = np.empty(self.nparams, np.float32)
        HH.check(_L.hn_gnn_get_params(self.h, p.ctypes.data_as(_fp)))
        return p

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.nparams
        HH.check(_L.hn_gnn_set_params(self.h, p.ctypes.data_as(_fp)))

    def forward(self, g: Graph, val, rhs, c):
        """Device arrays in (FP64 val/rhs/c), host (log x, log s, y) out."""
        lx = HH.DeviceArray(g.n, np.float32)
        ls = HH.DeviceArray(g.n, np.float32)
        ly = HH.DeviceArray(g.m, np.float32)
        HH.check(_L.hn_gnn_forward_dev(self.h, g.h, val.ptr, rhs.ptr, c.ptr, lx.ptr, ls.ptr, ly.ptr, None))
        HH.stream_sync()
        return lx.to_numpy(), ls.to_numpy(), ly.to_numpy()

    def warm_start(self, g: Graph, val, rhs, c):
        """(x0, y0, s0) on the host, built on the device (hn_gnn_warm_start_dev)."""
        x0 = HH.DeviceArray(g.n)
        y0 = HH.DeviceArray(g.m)
        s0 = HH.DeviceArray(g.n)
        HH.check(_L.hn_gnn_warm_start_dev(self.h, g.h, val.ptr, rhs.ptr, c.ptr, x0.ptr, y0.ptr, s0.ptr, None))
        HH.stream_sync()
        return x0.to_numpy(), y0.to_numpy(), s0.to_numpy()

    def gradients(self, g: Graph, val, rhs, c, tx, ts, ty):
        grad = np.empty(self.nparams, np.float32)
        loss = C.c_float()
        HH.check(_L.hn_gnn_gradients_dev(self.h, g.h, val.ptr, rhs.ptr, c.ptr, tx.ptr, ts.ptr, ty.ptr,
                                         grad.ctypes.data_as(_fp), C.byref(loss)))
        return float(loss.value), grad

    def train_step(self, g: Graph, val_ptr, rhs_ptr, c_ptr, tx_ptr, ts_ptr, ty_ptr, lr: float, loss_ptr) -> None:
        HH.check(_L.hn_gnn_train_step_dev(self.h, g.h, val_ptr, rhs_ptr, c_ptr, tx_ptr, ts_ptr, ty_ptr,
                                          C.c_float(lr), loss_ptr, None))


def _ptr_at(base: HH.DeviceArray, offset_elems: int) -> _vp:
    return _vp(int(base.ptr.value) + int(offset_elems) * int(base.dtype.itemsize))


@dataclass
class Dataset:
    """Training pairs resident in device memory: per instance the standard
    form's rhs and c (FP64) and the targets log x, log s, y (FP32); the
    pattern and values of M are shared by the family."""
    graph: Graph
    val: HH.DeviceArray
    rhs: HH.DeviceArray
    c: HH.DeviceArray
    tx: HH.DeviceArray
    ts: HH.DeviceArray
    ty: HH.DeviceArray
    count: int
    seeds: np.ndarray
    cold_iterations: np.ndarray


def make_dataset(make_lp, seeds: Sequence[int], chunk: int = 1024, pivot_rule: int = HH.HN_PIVOT_SKIP,
                 mu_target: float = MU_TARGET, log=print) -> Dataset:
    """Solves each seed's LP with the batched device IPM from the standard
    start and keeps the first iterate with mu <= mu_target as its target.
    All instances must share one standard-form pattern and values of M."""
    seeds = np.asarray(list(seeds), np.int64)
    sf0 = HH.lp_standard_form(make_lp(int(seeds[0])))
    m, n, nnz = sf0["m"], sf0["n"], sf0["ci"].size
    g = Graph(m, n, sf0["rp"], sf0["ci"])
    K = seeds.size
    rhs = HH.DeviceArray(K * m, np.float64)
    c = HH.DeviceArray(K * n, np.float64)
    tx = HH.DeviceArray(K * n, np.float32)
    ts = HH.DeviceArray(K * n, np.float32)
    ty = HH.DeviceArray(K * m, np.float32)
    its = np.zeros(K, np.int64)
    t0 = time.time()
    from concurrent.futures import ThreadPoolExecutor
    pool = ThreadPoolExecutor(max_workers=16)   # the ctypes calls release the GIL
    for a in range(0, K, chunk):
        idx = range(a, min(K, a + chunk))
        lps = list(pool.map(lambda i: make_lp(int(seeds[i])), idx))
        sfs = list(pool.map(HH.lp_standard_form, lps))
        res = HH.ipm_solve_batch(lps, want_points=True, pivot_rule=pivot_rule, mu_target=mu_target)
        R = np.empty((len(lps), m))
        Cc = np.empty((len(lps), n))
        X = np.empty((len(lps), n), np.float32)
        S = np.empty((len(lps), n), np.float32)
        Y = np.empty((len(lps), m), np.float32)
        for q, (sf, r) in enumerate(zip(sfs, res)):
            if sf["m"] != m or sf["n"] != n or sf["ci"].size != nnz or not np.array_equal(sf["ci"], sf0["ci"]) \
                    or not np.array_equal(sf["v"], sf0["v"]):
                raise ValueError("make_dataset: instances of one family must share the matrix M")
            R[q], Cc[q] = sf["rhs"], sf["c"]
            X[q], S[q], Y[q] = np.log(r.x), np.log(r.s), r.y
            its[a + q] = r.iterations
        rhs_h = R.ravel()
        C_h = Cc.ravel()
        for dev, host, off in ((rhs, rhs_h, a * m), (c, C_h, a * n), (tx, X.ravel(), a * n), (ts, S.ravel(), a * n),
                               (ty, Y.ravel(), a * m)):
            HH.check(HH.LIB.hn_memcpy_h2d(_ptr_at(dev, off), host.ctypes.data_as(_vp), host.nbytes))
        log(f"[warmstart] targets {a + len(lps)}/{K} ({time.time() - t0:.1f}s)")
    pool.shutdown()
    val = HH.DeviceArray.from_numpy(sf0["v"])
    return Dataset(g, val, rhs, c, tx, ts, ty, K, seeds, its)


def train(net: Network, ds: Dataset, epochs: int, lr: float = 1e-3, lr_final: float = 1e-4, seed: int = 0,
          log=print) -> list:
    """Adam over the instances in a seeded random order per epoch, learning
    rate decayed geometrically from lr to lr_final.  Returns per-epoch mean
    losses (read back once per epoch)."""
    m, n = ds.graph.m, ds.graph.n
    rng = np.random.default_rng(seed)
    steps = epochs * ds.count
    losses = HH.DeviceArray(ds.count, np.float32)
    hist = []
    k = 0
    t0 = time.time()
    for ep in range(epochs):
        for i in rng.permutation(ds.count):
            frac = k / max(1, steps - 1)
            step_lr = lr * (lr_final / lr) ** frac
            net.train_step(ds.graph, ds.val.ptr, _ptr_at(ds.rhs, i * m), _ptr_at(ds.c, i * n), _ptr_at(ds.tx, i * n),
                           _ptr_at(ds.ts, i * n), _ptr_at(ds.ty, i * m), step_lr, _ptr_at(losses, int(i)))
            k += 1
        HH.stream_sync()
        ep_loss = float(losses.to_numpy().mean())
        hist.append(ep_loss)
        log(f"[warmstart] epoch {ep}: mean loss {ep_loss:.4f} ({time.time() - t0:.1f}s)")
    return hist


def _same_pattern(sf, m: int, n: int, row_ptr, col) -> bool:
    return sf["m"] == m and sf["n"] == n and sf["ci"].size == col.size and \
        np.array_equal(sf["rp"], row_ptr) and np.array_equal(sf["ci"], col)


def predict(net: Network, lps, graph: Optional[Graph] = None, chunk: int = 256):
    """Warm starts (x0, y0, s0) of the scaled standard forms of `lps`:
    x0 = exp(log x), y0 = y, s0 = max(c - M'y, exp(log s)), formed on the
    device.  The dual-feasible s0 matters: a start whose mu is small next to
    its dual residual lets mu collapse first and the dual steps then stall.
    Runs of up to `chunk` instances with one sparsity pattern share one
    upload, back-to-back network launches on the library stream and one
    download."""
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=16) as pool:   # the ctypes calls release the GIL
        sfs = list(pool.map(HH.lp_standard_form, lps))
    out = [None] * len(sfs)
    k = 0
    while k < len(sfs):
        sf = sfs[k]
        if graph is None or not _same_pattern(sf, graph.m, graph.n, graph.row_ptr, graph.col):
            graph = Graph(sf["m"], sf["n"], sf["rp"], sf["ci"])  # kept for the next instances of this pattern
        j = k + 1
        while j < len(sfs) and j - k < chunk and _same_pattern(sfs[j], graph.m, graph.n, graph.row_ptr, graph.col):
            j += 1
        cnt, m, n, nnz = j - k, graph.m, graph.n, graph.nnz
        val = HH.DeviceArray.from_numpy(np.concatenate([q["v"] for q in sfs[k:j]]))
        rhs = HH.DeviceArray.from_numpy(np.concatenate([q["rhs"] for q in sfs[k:j]]))
        c = HH.DeviceArray.from_numpy(np.concatenate([q["c"] for q in sfs[k:j]]))
        x0, y0, s0 = HH.DeviceArray(cnt * n), HH.DeviceArray(cnt * m), HH.DeviceArray(cnt * n)
        for q in range(cnt):
            HH.check(_L.hn_gnn_warm_start_dev(net.h, graph.h, _ptr_at(val, q * nnz), _ptr_at(rhs, q * m),
                                              _ptr_at(c, q * n), _ptr_at(x0, q * n), _ptr_at(y0, q * m),
                                              _ptr_at(s0, q * n), None))
        HH.stream_sync()
        X, Y, S = x0.to_numpy().reshape(cnt, n), y0.to_numpy().reshape(cnt, m), s0.to_numpy().reshape(cnt, n)
        for q in range(cnt):
            out[k + q] = (X[q], Y[q], S[q])
        k = j
    return out


def evaluate(net: Network, make_lp, seeds: Sequence[int], graph: Optional[Graph] = None,
             pivot_rule: int = HH.HN_PIVOT_SKIP, log=print, gap_tol: float = 0.0, on_device: bool = True,
             endgame_mu: float = 0.0) -> dict:
    """Cold vs warm-started batched solves of the held-out seeds: iterations,
    wall time (the warm time includes the network's inference), acceptance
    of the predictions, and agreement of the two optima.  on_device: the
    network runs inside the batched call on the packed standard forms
    (hn_ipm_solve_batch_gnn); else predict() runs first and the points are
    passed in.  gap_tol > 0: both solves also stop only at a relative gap <=
    gap_tol (see hn_ipm_opts).  endgame_mu > 0: both solves may end by the
    optimal-face projection (hn_ipm_opts.endgame_mu)."""
    lps = [make_lp(int(s)) for s in seeds]
    if graph is None:
        sf = HH.lp_standard_form(lps[0])
        graph = Graph(sf["m"], sf["n"], sf["rp"], sf["ci"])
    kw = dict(pivot_rule=pivot_rule, gap_tol=gap_tol, endgame_mu=endgame_mu)
    # warm-up at full size (one iteration), so that both timed calls reuse
    # the grown device arena and page-locked staging
    HH.ipm_solve_batch(lps, max_iterations=1, **kw)
    if on_device:
        HH.ipm_solve_batch(lps[:4], gnn=(net.h, graph.h), **kw)
    t0 = time.perf_counter()
    cold = HH.ipm_solve_batch(lps, **kw)
    t_cold = time.perf_counter() - t0
    t_pred = 0.0
    if on_device:
        t0 = time.perf_counter()
        warm = HH.ipm_solve_batch(lps, gnn=(net.h, graph.h), **kw)
        t_warm = time.perf_counter() - t0
    else:
        t0 = time.perf_counter()
        warm_pts = predict(net, lps, graph)
        t_pred = time.perf_counter() - t0
        t0 = time.perf_counter()
        warm = HH.ipm_solve_batch(lps, warm=warm_pts, **kw)
        t_warm = time.perf_counter() - t0
    ci = np.array([r.iterations for r in cold])
    wi = np.array([r.iterations for r in warm])
    acc = np.array([r.warm_started for r in warm])
    rel = np.array([abs(a.objective_value - b.objective_value) / max(1.0, abs(b.objective_value))
                    for a, b in zip(warm, cold)])
    out = {
        "instances": len(lps),
        "cold_iterations_mean": float(ci.mean()), "warm_iterations_mean": float(wi.mean()),
        "cold_iterations_max": int(ci.max()), "warm_iterations_max": int(wi.max()),
        "cold_iterations_median": float(np.median(ci)), "warm_iterations_median": float(np.median(wi)),
        "cold_stalled": int((ci >= 100).sum()), "warm_stalled": int((wi >= 100).sum()),
        "accepted": int(acc.sum()),
        "cold_optimal": int(sum(r.status == "OPTIMAL" for r in cold)),
        "warm_optimal": int(sum(r.status == "OPTIMAL" for r in warm)),
        "max_rel_objective_diff": float(rel.max()),
        "cold_wall_s": round(t_cold, 4), "warm_wall_s": round(t_warm + t_pred, 4),
        "inference": "on the device inside the batched call" if on_device else "predict() before the call",
        "predict_s": round(t_pred, 4), "warm_solve_s": round(t_warm, 4),
        "iteration_ratio": float(ci.mean() / max(wi.mean(), 1e-9)),
        "gap_tol": gap_tol,
        "cold_max_rel_gap": float(max(r.rel_gap for r in cold)),
        "warm_max_rel_gap": float(max(r.rel_gap for r in warm)),
        "cold_max_rel_primal": float(max(r.rel_primal for r in cold)),
        "warm_max_rel_primal": float(max(r.rel_primal for r in warm)),
        "endgame_mu": endgame_mu,
        "cold_projected": int(sum(r.projected for r in cold)),
        "warm_projected": int(sum(r.projected for r in warm)),
        "cold_kkt_1e-8": int(sum(max(r.rel_primal, r.rel_dual, r.rel_gap) <= 1e-8 for r in cold)),
        "warm_kkt_1e-8": int(sum(max(r.rel_primal, r.rel_dual, r.rel_gap) <= 1e-8 for r in warm)),
    }
    log(f"[warmstart] eval {out}")
    return out


def transport_family(S: int = 200, T: int = 300):
    import lpgen
    return lambda seed: lpgen.transport_lp(seed, S, T)


def run_experiment(n_train: int = 20000, n_test: int = 1000, epochs: int = 6, S: int = 200, T: int = 300,
                   lr: float = 1e-3, lr_final: float = 1e-4, out_params: Optional[str] = None, seed: int = 0,
                   log=print) -> dict:
    """The transportation warm-start workload: train on seeds 0..n_train-1,
    test on seeds 1,000,000.. (held out), report iterations and wall time to
    the 1e-8 stopping test against cold starts."""
    HH.init(-1)
    make = transport_family(S, T)
    t0 = time.time()
    ds = make_dataset(make, range(n_train), log=log)
    t_data = time.time() - t0
    net = Network(seed)
    t0 = time.time()
    hist = train(net, ds, epochs, lr=lr, lr_final=lr_final, seed=seed, log=log)
    t_train = time.time() - t0
    if out_params:
        np.save(out_params, net.get_params())
    res = evaluate(net, make, range(1_000_000, 1_000_000 + n_test), ds.graph, log=log)
    res.update({"train_instances": n_train, "epochs": epochs, "loss_history": [round(v, 5) for v in hist],
                "target_iterations_mean": float(ds.cold_iterations.mean()),
                "data_s": round(t_data, 2), "train_s": round(t_train, 2), "family": f"transport_{S}x{T}"})
    return res


if __name__ == "__main__":
    import argparse
    import json
    ap = argparse.ArgumentParser(description="train and evaluate the transportation warm start")
    ap.add_argument("--train", type=int, default=20000)
    ap.add_argument("--test", type=int, default=1000)
    ap.add_argument("--epochs", type=int, default=6)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out-params", default=None)
    ap.add_argument("--report", default=None)
    ap.add_argument("--params", default=None, help="evaluate these trained parameters (.npy) without training")
    ap.add_argument("--gap-tol", type=float, default=0.0, help="with --params: relative-gap bar of both solves")
    ap.add_argument("--endgame-mu", type=float, default=0.0, help="with --params: optimal-face projection from this mu")
    a = ap.parse_args()
    if a.params:
        HH.init(-1)
        net = Network(0)
        net.set_params(np.load(a.params))
        r = evaluate(net, transport_family(), range(1_000_000, 1_000_000 + a.test), log=lambda s: print(s, flush=True),
                     gap_tol=a.gap_tol, endgame_mu=a.endgame_mu)
    else:
        r = run_experiment(a.train, a.test, a.epochs, lr=a.lr, out_params=a.out_params,
                           log=lambda s: print(s, flush=True))
    print(json.dumps(r), flush=True)
    if a.report:
        with open(a.report, "w") as f:
            json.dump(r, f, indent=1)
