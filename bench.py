#!/usr/bin/env python3
"""Benchmark of the MI355X LP hot path.

Default workload (one GPU): the SpMV microbenchmark - CSR 2,000,000 x
5,000,000 with 40M nonzeros and power-law (Zipf s=1.2, cap 10,000) row
lengths, seed 2.  One step = y = A x and z = A' w in FP64 (A' is the
pre-built transpose, so both are gathers).  Inputs are resident in HBM before
the timed region; the timed region is bracketed by stream synchronisations
and measured with HIP events on the stream the kernels run on.

Printed: one JSON line with the metric, the roofline of the dominant kernel
(the FP64 A x product) and the reference CPU baseline timed on this host.

The LP legs ("lp" object, skipped by --no-lp) report LP solves/s:
  * batch: 4,096 LPs (instance seeds 1000..5095, m log-uniform in [50, 2000],
    n = 2.5 m, 5 nonzeros per column) solved by one persistent launch, under
    the reference's pivot rule and under the skip rule, against the plain
    FP64 single-LP GPU solver (reference rule) run sequentially on a strided
    sample of the same instances (sizes are i.i.d., so the sample is unbiased);
  * single: the 2,000 x 5,000 random LP (seed 1), plain FP64 vs mixed, under
    both pivot rules;
  * cpu_baseline: the reference InteriorPointSolver (crossover off, 1 thread)
    on a prefix of the same strided sample, under a per-instance time limit.
LP timings are host wall-clock around the whole call (standard form on the
host, upload, solve, download), so they are end-to-end, not HBM-resident.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "indigenous-cpu-gpu-optimization-engine-for-indian-industry_amd")
sys.path.insert(0, PKG)

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` as measured by the rocprofv3
    FETCH_SIZE / WRITE_SIZE runs summarised in profiles/ (scaled by the copy
    kernel calibration written there); None when no summary covers it."""
    d = os.path.join(ROOT, "profiles")
    if not os.path.isdir(d):
        return None
    for fn in sorted(os.listdir(d)):
        if not fn.endswith("_pmc.md"):
            continue
        for line in open(os.path.join(d, fn)):
            parts = line.split()
            if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == kernel:
                return int(parts[2])
    return None


def host_info() -> dict:
    """The host the CPU baselines ran on: usable cores and the CPU model."""
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 1
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": usable, "cpu_count": os.cpu_count(), "cpu_model": model}


def all_core_threads() -> int:
    # the GPU box gives one GPU's job a 16-core share of its host
    return max(1, min(16, host_info()["nproc"]))


def spmv_bytes(nrows: int, ncols: int, nnz: int, sv: int) -> int:
    """Algorithmic bytes of one CSR product: values + 4-byte column indices,
    4-byte row pointers, x once and y once."""
    return nnz * (sv + 4) + (nrows + 1) * 4 + ncols * sv + nrows * sv


def cpu_baseline(case, target_s: float) -> dict:
    """The reference's own CPU SpMV (CPUBackend::spmv and the row-scatter
    transpose product) on 1 thread, on a bounded number of full products."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # checker / baseline only
    ab = spmv_bytes(case.nrows, case.ncols, case.nnz, 8)
    tb = spmv_bytes(case.ncols, case.nrows, case.nnz, 8)
    if O.have_reference():
        t1 = O.ref_cpu_spmv_seconds(case.row_ptr, case.col, case.val, case.x, 1)
        reps = max(1, int(target_s / 2 / max(t1, 1e-3)))
        ta = O.ref_cpu_spmv_seconds(case.row_ptr, case.col, case.val, case.x, reps) / reps
        tt = O.ref_cpu_spmv_t_seconds(case.row_ptr, case.col, case.val, case.w, case.ncols, reps) / reps
        kind = "reference"
        sample = (f"{reps} full FP64 products each of A x (CPUBackend::spmv) and A' w "
                  f"(SparseMatrix::transpose_multiply) on the 2M x 5M, 40M-nnz matrix")
    else:
        t1 = O.port_spmv_seconds(case.row_ptr, case.col, case.val, case.x, 1)
        reps = max(1, int(target_s / 2 / max(t1, 1e-3)))
        ta = O.port_spmv_seconds(case.row_ptr, case.col, case.val, case.x, reps) / reps
        tt = ta
        kind = "port"
        sample = f"{reps} full FP64 A x products (restated CPU loop) on the 2M x 5M matrix"
    return {"value": round((ab + tb) / (ta + tt) / 1e9, 3), "unit": "GB/s", "cores": 1, "kind": kind,
            "sample": sample, "ms_per_step": round(1e3 * (ta + tt), 3), "host": host_info()}


def cpu_all_core_spmv(case, target_s: float) -> dict:
    """This repository's own CPU loop (sequential sum per row, as the
    reference's) with the rows split over every usable host core, on A and on
    the host transpose.  Labelled: not the reference's kernel."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # baseline only
    T = all_core_threads()
    trp, tci, tv = O.csr_transpose(case.row_ptr, case.col, case.val, case.ncols)
    t1 = O.port_spmv_mt_seconds(case.row_ptr, case.col, case.val, case.x, 1, T)
    reps = max(1, int(target_s / 2 / max(t1, 1e-3)))
    ta = O.port_spmv_mt_seconds(case.row_ptr, case.col, case.val, case.x, reps, T) / reps
    tt = O.port_spmv_mt_seconds(trp, tci, tv, case.w, reps, T) / reps
    ab = spmv_bytes(case.nrows, case.ncols, case.nnz, 8)
    tb = spmv_bytes(case.ncols, case.nrows, case.nnz, 8)
    return {"value": round((ab + tb) / (ta + tt) / 1e9, 3), "unit": "GB/s", "cores": T, "kind": "port",
            "sample": (f"{reps} FP64 products each of A x and A' w (host CSR transpose) on the 2M x 5M matrix, "
                       f"rows split over {T} threads by nonzero count; this repository's loop, not the "
                       f"reference's kernel"),
            "ms_per_step": round(1e3 * (ta + tt), 3), "host": host_info()}


def run_spmv(args) -> dict:
    import hypernova_hip as H
    import lpgen

    t0 = time.time()
    case = lpgen.powerlaw_spmv(seed=2)
    log(f"generated {case.nrows}x{case.ncols} nnz={case.nnz} in {time.time() - t0:.1f}s")
    info = H.init(-1)
    log(f"device {info.name.decode()} {info.arch.decode()} CUs={info.compute_units}")
    A = H.CsrMatrix.upload(case.val, case.col, case.row_ptr, case.ncols, H.HN_F64 | H.HN_F32)
    At = A.transpose()
    log("uploaded A and built A' on the device")
    # layouts: autotuned, or fixed by --chunks-A / --chunks-At (so that a
    # profiler run, whose counters distort the autotune's timings, measures
    # the layout the plain run chose)
    tA = tT = float("nan")
    if args.chunks_A:
        A.set_chunks(*(int(t) for t in args.chunks_A.split(",")))
    else:
        tA = A.autotune(10)
    if args.chunks_At:
        At.set_chunks(*(int(t) for t in args.chunks_At.split(",")))
    else:
        tT = At.autotune(10)
    log(f"layouts: A tiling {A.tiling()} slab {A.slab()} chunks {A.chunks()} ({tA:.1f} us), "
        f"A' tiling {At.tiling()} slab {At.slab()} chunks {At.chunks()} ({tT:.1f} us)")

    x = H.DeviceArray.from_numpy(case.x)
    w = H.DeviceArray.from_numpy(case.w)
    y = H.DeviceArray(case.nrows)
    z = H.DeviceArray(case.ncols)
    x32 = H.DeviceArray.from_numpy(case.x.astype(np.float32))
    w32 = H.DeviceArray.from_numpy(case.w.astype(np.float32))
    y32 = H.DeviceArray(case.nrows, np.float32)
    z32 = H.DeviceArray(case.ncols, np.float32)

    def step64():
        A.spmv_dev(x, y, H.HN_F64)
        At.spmv_dev(w, z, H.HN_F64)

    for _ in range(args.warmup):
        step64()
    H.stream_sync()
    e0, e1 = H.Event(), H.Event()
    # timed region: K steps
    H.stream_sync()
    e0.record()
    for _ in range(args.steps):
        step64()
    e1.record()
    H.stream_sync()
    ms_step = e0.elapsed_ms(e1) / args.steps

    # dominant kernel alone: FP64 A x, averaged over K launches
    reps = max(args.steps, 20)
    e0.record()
    for _ in range(reps):
        A.spmv_dev(x, y, H.HN_F64)
    e1.record()
    H.stream_sync()
    ms_ax = e0.elapsed_ms(e1) / reps
    e0.record()
    for _ in range(reps):
        At.spmv_dev(w, z, H.HN_F64)
    e1.record()
    H.stream_sync()
    ms_atw = e0.elapsed_ms(e1) / reps
    # FP32 variants (reported alongside)
    for _ in range(3):
        A.spmv_dev(x32, y32, H.HN_F32)
        At.spmv_dev(w32, z32, H.HN_F32)
    e0.record()
    for _ in range(reps):
        A.spmv_dev(x32, y32, H.HN_F32)
    e1.record()
    H.stream_sync()
    ms_ax32 = e0.elapsed_ms(e1) / reps
    e0.record()
    for _ in range(reps):
        At.spmv_dev(w32, z32, H.HN_F32)
    e1.record()
    H.stream_sync()
    ms_atw32 = e0.elapsed_ms(e1) / reps

    # copy kernel of known size (hardware-counter calibration reference)
    cb = 256 << 20
    c_src, c_dst = H.DeviceArray(cb // 8), H.DeviceArray(cb // 8)
    H.check(H.LIB.hn_copy_dev(c_dst.ptr, c_src.ptr, cb, None))
    e0.record()
    for _ in range(10):
        H.check(H.LIB.hn_copy_dev(c_dst.ptr, c_src.ptr, cb, None))
    e1.record()
    H.stream_sync()
    copy_gbs = 2 * cb * 10 / (e0.elapsed_ms(e1) * 1e-3) / 1e9

    # error against the CPU FP64 result (checker only)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    y_ref = O.spmv(case.row_ptr, case.col, case.val, case.x)
    z_ref = O.spmv_t(case.row_ptr, case.col, case.val, case.w, case.ncols)
    err_y = float(np.max(np.abs(y.to_numpy() - y_ref) / np.maximum(1.0, np.abs(y_ref))))
    err_z = float(np.max(np.abs(z.to_numpy() - z_ref) / np.maximum(1.0, np.abs(z_ref))))
    err_y32 = float(np.max(np.abs(y32.to_numpy() - y_ref) / np.maximum(1.0, np.abs(y_ref))))
    log(f"errors vs CPU FP64: Ax {err_y:.2e} A'w {err_z:.2e} Ax(fp32) {err_y32:.2e}")

    ab64 = spmv_bytes(case.nrows, case.ncols, case.nnz, 8)
    tb64 = spmv_bytes(case.ncols, case.nrows, case.nnz, 8)
    ab32 = spmv_bytes(case.nrows, case.ncols, case.nnz, 4)
    tb32 = spmv_bytes(case.ncols, case.nrows, case.nnz, 4)
    value = (ab64 + tb64) / (ms_step * 1e-3) / 1e9
    achieved = ab64 / (ms_ax * 1e-3) / 1e9
    if A.chunks()[0]:
        ci, sp, nt = A.chunks()
        kname = f"k_spmv_chunk<double,{ci},{nt}>"
        kdesc = f"{kname} (one A x product, chunked stream, chunks spanning <= 2^{sp} columns)"
    elif A.slab():
        kname = "k_spmv_slab<double,4>"
        kdesc = f"k_spmv_slab<double,4> (one A x product, slab layout 2^{A.slab()} columns)"
    else:
        kname = f"k_spmv_merge<double,{A.tiling()}>"
        kdesc = "k_spmv_merge<double> + k_spmv_fixup<double> (one A x product)"
    out = {
        "metric": "SpMV GB/s (y=A x and z=A' w, FP64 CSR, algorithmic bytes)",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator, seed 2)",
        "config": {"workload": "spmv_powerlaw_2Mx5M_40Mnnz_zipf1.2", "nrows": case.nrows, "ncols": case.ncols,
                   "nnz": case.nnz, "index_bytes": 4, "tiling_A": A.tiling(), "tiling_At": At.tiling(),
                   "slab_log2_A": A.slab(), "slab_log2_At": At.slab(),
                   "chunks_A": list(A.chunks()), "chunks_At": list(At.chunks())},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(kname),
                     "kernel": kdesc,
                     "us_per_launch": round(ms_ax * 1e3, 2), "bytes_per_launch": ab64},
        "detail": {
            "ax_f64_us": round(ms_ax * 1e3, 2), "ax_f64_gbs": round(ab64 / (ms_ax * 1e-3) / 1e9, 1),
            "atw_f64_us": round(ms_atw * 1e3, 2), "atw_f64_gbs": round(tb64 / (ms_atw * 1e-3) / 1e9, 1),
            "ax_f32_us": round(ms_ax32 * 1e3, 2), "ax_f32_gbs": round(ab32 / (ms_ax32 * 1e-3) / 1e9, 1),
            "atw_f32_us": round(ms_atw32 * 1e3, 2), "atw_f32_gbs": round(tb32 / (ms_atw32 * 1e-3) / 1e9, 1),
            "err_ax_f64": err_y, "err_atw_f64": err_z, "err_ax_f32": err_y32,
            "copy_kernel_gbs": round(copy_gbs, 1),
        },
    }
    if not args.no_cpu_baseline:
        t = time.time()
        out["cpu_baseline"] = cpu_baseline(case, args.cpu_seconds)
        log(f"cpu baseline {out['cpu_baseline']} in {time.time() - t:.1f}s")
        t = time.time()
        out["cpu_baseline_all_cores"] = cpu_all_core_spmv(case, args.cpu_seconds / 2)
        log(f"cpu all-core baseline {out['cpu_baseline_all_cores']} in {time.time() - t:.1f}s")
    return out


# The LP configs' bar: the reference's absolute stopping test (interior_point.cpp
# :525-530) and, on top, relative KKT residuals and gap <= 1e-8.  Every LP leg
# below solves with gap_tol = KKT_TOL, so the plain and the chosen solvers are
# compared at the same bar.
KKT_TOL = 1e-8


def _kkt_max(results) -> dict:
    return {"max_rel_primal": max(r.rel_primal for r in results),
            "max_rel_dual": max(r.rel_dual for r in results),
            "max_rel_gap": max(r.rel_gap for r in results),
            "kkt_1e-8": sum(r.rel_primal <= KKT_TOL and r.rel_dual <= KKT_TOL and r.rel_gap <= KKT_TOL
                            for r in results)}


def run_lp(args) -> dict:
    import hypernova_hip as H
    import lpgen

    H.init(-1)
    out = {}
    # batched instances: one persistent launch
    count = args.batch
    t0 = time.time()
    lps = [lpgen.batch_lp(i) for i in range(count)]
    log(f"generated {count} batch LPs (sum m={sum(l.m for l in lps)}) in {time.time() - t0:.1f}s")
    # warm-up: code objects, device arena and page-locked staging sized for this batch
    H.ipm_solve_batch(lps, pivot_rule=H.HN_PIVOT_SKIP, gap_tol=KKT_TOL)
    batch = {"workload": f"batch_{count}_lps_m50-2000_n2.5m_5nnz_per_col", "lps": count,
             "stop": "reference test (absolute 1e-8) and relative gap <= 1e-8"}
    runs = {}
    for rule, rname in ((H.HN_PIVOT_SKIP, "skip_rule"), (H.HN_PIVOT_CLAMP, "reference_rule")):
        t0 = time.perf_counter()
        res = H.ipm_solve_batch(lps, precision=H.HN_PREC_F64, pivot_rule=rule, verbose=True, gap_tol=KKT_TOL)
        wall = time.perf_counter() - t0
        status = {}
        for r in res:
            status[r.status] = status.get(r.status, 0) + 1
        its = [r.iterations for r in res]
        b = {"wall_s": round(wall, 3), "lps_per_s": round(count / wall, 2), "status": status,
             "iterations_mean": round(float(np.mean(its)), 2), "iterations_max": int(max(its)),
             "precision": "f64",
             "stop_test_1e-8": sum(r.rp_norm < 1e-8 and r.rd_norm < 1e-8 and r.mu < 1e-8 for r in res),
             "at_iteration_limit": sum(r.iterations >= 100 for r in res)}
        b.update(_kkt_max(res))
        batch[rname] = b
        runs[rname] = res
        log(f"batch {rname}: {count} LPs in {wall:.2f}s -> {count / wall:.1f} LPs/s {status}")
    # mixed precision in the batch: FP32 factor of N with FP64 refinement of
    # every Newton step, FP64 once mu < mu_switch, same stop bar
    for rule, rname in ((H.HN_PIVOT_SKIP, "skip_rule_mixed"),):
        t0 = time.perf_counter()
        res = H.ipm_solve_batch(lps, precision=H.HN_PREC_MIXED, pivot_rule=rule, gap_tol=KKT_TOL)
        wall = time.perf_counter() - t0
        status = {}
        for r in res:
            status[r.status] = status.get(r.status, 0) + 1
        its = [r.iterations for r in res]
        b = {"wall_s": round(wall, 3), "lps_per_s": round(count / wall, 2), "status": status,
             "iterations_mean": round(float(np.mean(its)), 2), "iterations_max": int(max(its)),
             "precision": "mixed (fp32 factor, fp64 refinement)",
             "fp64_from_iter_mean": round(float(np.mean([r.fp64_from_iter for r in res])), 2),
             "at_iteration_limit": sum(r.iterations >= 100 for r in res)}
        b.update(_kkt_max(res))
        batch[rname] = b
        log(f"batch {rname}: {count} LPs in {wall:.2f}s -> {count / wall:.1f} LPs/s {status}")
    batch["mixed_speedup_vs_f64_batch"] = round(batch["skip_rule_mixed"]["lps_per_s"] /
                                                batch["skip_rule"]["lps_per_s"], 3)
    # autotuned precision: per m bucket, FP64 or the FP32 factor, whichever
    # solved a sample of the bucket faster (a one-time cost, timed apart)
    sample = lps[::max(1, count // 512)]
    t0 = time.perf_counter()
    table = H.batch_autotune(sample, per_bucket=args.tune_per_bucket, pivot_rule=H.HN_PIVOT_SKIP, gap_tol=KKT_TOL)
    tune_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    res = H.ipm_solve_batch(lps, precision=H.HN_PREC_AUTO, pivot_rule=H.HN_PIVOT_SKIP, gap_tol=KKT_TOL)
    wall = time.perf_counter() - t0
    status = {}
    for r in res:
        status[r.status] = status.get(r.status, 0) + 1
    b = {"wall_s": round(wall, 3), "lps_per_s": round(count / wall, 2), "status": status,
         "iterations_mean": round(float(np.mean([r.iterations for r in res])), 2),
         "precision": "per m bucket from hn_batch_autotune", "autotune_s": round(tune_s, 2),
         "autotune_sample": f"{len(sample)} LPs, at most {args.tune_per_bucket} per bucket",
         "table": table, "at_iteration_limit": sum(r.iterations >= 100 for r in res)}
    b.update(_kkt_max(res))
    batch["skip_rule_autotuned"] = b
    log(f"batch skip_rule_autotuned: {count} LPs in {wall:.2f}s -> {count / wall:.1f} LPs/s {status} "
        f"(autotune {tune_s:.1f}s: {[(t['m_hi'], t['precision']) for t in table]})")
    res = runs["reference_rule"]
    # plain FP64 single-LP GPU solver, sequential, on a strided sample
    stride = max(1, count // args.seq_sample)
    idx = list(range(0, count, stride))[:args.seq_sample]
    H.ipm_solve(lps[idx[0]], gap_tol=KKT_TOL)
    t0 = time.perf_counter()
    seq = [H.ipm_solve(lps[i], gap_tol=KKT_TOL) for i in idx]
    seq_wall = time.perf_counter() - t0
    seq_rate = len(idx) / seq_wall
    seq_status = {}
    for r in seq:
        seq_status[r.status] = seq_status.get(r.status, 0) + 1
    batch["sequential_plain"] = {"lps": len(idx), "stride": stride, "wall_s": round(seq_wall, 3),
                                 "lps_per_s": round(seq_rate, 2), "status": seq_status,
                                 "same_status_as_batch": sum(r.status == res[i].status for r, i in zip(seq, idx))}
    batch["sequential_plain"].update(_kkt_max(seq))
    batch["speedup_vs_sequential_plain"] = {
        k: round(batch[k]["lps_per_s"] / seq_rate, 2) for k in ("skip_rule", "reference_rule")}
    log(f"sequential plain: {len(idx)} LPs in {seq_wall:.2f}s -> {seq_rate:.2f} LPs/s")
    out["batch"] = batch

    # one 2,000 x 5,000 LP, plain FP64 vs mixed precision
    lp = lpgen.random_lp(1, 2000, 5000)
    H.ipm_solve(lp, gap_tol=KKT_TOL)
    single = {"workload": "random_lp_2000x5000_5nnz_per_col_seed1",
              "stop": "reference test (absolute 1e-8) and relative gap <= 1e-8"}
    for name, prec, rule in (("plain_f64", H.HN_PREC_F64, H.HN_PIVOT_CLAMP), ("mixed", H.HN_PREC_MIXED, H.HN_PIVOT_CLAMP),
                             ("skip_rule_f64", H.HN_PREC_F64, H.HN_PIVOT_SKIP),
                             ("skip_rule_mixed", H.HN_PREC_MIXED, H.HN_PIVOT_SKIP)):
        t0 = time.perf_counter()
        r = H.ipm_solve(lp, precision=prec, pivot_rule=rule, gap_tol=KKT_TOL)
        dt = time.perf_counter() - t0
        single[name] = {"ms": round(1e3 * dt, 2), "status": r.status, "iterations": r.iterations,
                        "objective": r.objective_value, "rel_primal": r.rel_primal, "rel_dual": r.rel_dual,
                        "rel_gap": r.rel_gap, "fp64_from_iter": r.fp64_from_iter}
        log(f"single {name}: {r.status} it={r.iterations} {1e3 * dt:.1f} ms")
    out["single"] = single

    # the large power-law LP's generator (column counts on [2, 200], rows
    # scaled by 10^U(-2,2), seed 3): at reduced sizes, which the dense normal
    # equations solve to the 1e-8 bar, plain FP64 against mixed precision;
    # at the full 200,000 x 500,000 size, where only the matrix-free
    # Jacobi-PCG path exists, what plain and mixed reach under a time limit
    out["large"] = run_large(args, H, lpgen)

    # learned warm start: the 200 x 300 transportation family, the committed
    # trained network on 1,000 held-out seeds, cold vs warm batched solves
    import warmstart as WS
    net = WS.Network(0)
    net.set_params(np.load(os.path.join(PKG, "data", "warmstart_transport_200x300.npy")))
    # This degenerate family (rank m-1) stalls in its dual steps once mu is
    # near 1e-8 and does not reach a relative gap of 1e-8 within 100
    # iterations, so both solves stop at the reference's test; the relative
    # gaps they reach are reported.
    ev = WS.evaluate(net, WS.transport_family(), range(1_000_000, 1_000_000 + args.ws_test),
                     log=lambda *a: None)
    ev["workload"] = f"transport_200x300_heldout_{args.ws_test}"
    ev["stop"] = "reference test (absolute 1e-8 on rp, rd, mu)"
    # to the 1e-8 relative-gap bar: the optimal-face projection end game
    # (opt-in, not in the reference) finishes the stalled dual steps
    ev8 = WS.evaluate(net, WS.transport_family(), range(1_000_000, 1_000_000 + args.ws_test),
                      log=lambda *a: None, gap_tol=KKT_TOL, endgame_mu=1e-6)
    ev8["stop"] = "reference test and relative gap <= 1e-8, optimal-face projection from mu <= 1e-6"
    ev["to_1e-8_gap"] = ev8
    log(f"warm start to 1e-8 gap: iterations {ev8['cold_iterations_mean']:.2f} cold -> "
        f"{ev8['warm_iterations_mean']:.2f} warm, wall {ev8['cold_wall_s']:.2f} s -> {ev8['warm_wall_s']:.2f} s, "
        f"kkt {ev8['cold_kkt_1e-8']}/{ev8['warm_kkt_1e-8']}")
    out["warm_start"] = ev
    log(f"warm start: iterations {ev['cold_iterations_mean']:.2f} cold -> {ev['warm_iterations_mean']:.2f} warm, "
        f"wall {ev['cold_wall_s']:.2f} s -> {ev['warm_wall_s']:.2f} s (inference {ev['inference']})")

    # Netlib: the reference's 21 models (tests/golden/netlib/), one GPU solve
    # each under the reference's rules, against HiGHS's optimum
    import mps
    fx = json.load(open(os.path.join(ROOT, "tests", "golden", "netlib.json")))
    models = {name: mps.read_mps(os.path.join(ROOT, "tests", "golden", "netlib", name + ".mps.gz"))[0]
              for name in sorted(fx)}
    H.ipm_solve(models["afiro"])
    per, t_all = {}, 0.0
    for name, mlp in models.items():
        t0 = time.perf_counter()
        r = H.ipm_solve(mlp)
        dt = time.perf_counter() - t0
        t_all += dt
        per[name] = {"status": r.status, "iterations": r.iterations, "ms": round(1e3 * dt, 2),
                     "rel_err_vs_highs": abs(r.objective_value - fx[name]["highs"]) / max(1.0, abs(fx[name]["highs"])),
                     "reference_ipm": f"{fx[name]['ref_ipm_status']}/{fx[name]['ref_ipm_iterations']}"}
    out["netlib"] = {"models": len(models), "wall_s": round(t_all, 3),
                     "within_1e-6_of_highs": sum(v["rel_err_vs_highs"] <= 1e-6 for v in per.values()),
                     "same_status_and_iterations_as_reference_ipm": sum(
                         f"{v['status']}/{v['iterations']}" == v["reference_ipm"] for v in per.values()),
                     "per_model": per}
    log(f"netlib: 21 models in {t_all:.2f} s, {out['netlib']['within_1e-6_of_highs']} within 1e-6 of HiGHS, "
        f"{out['netlib']['same_status_and_iterations_as_reference_ipm']} with the reference IPM's status/iterations; "
        f"dfl001 {per['dfl001']['status']} in {per['dfl001']['ms']:.0f} ms")

    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O  # baseline only
        if O.have_reference():
            solve, kind = O.ref_ipm_solve, "reference"
        else:
            solve, kind = O.ipm_solve, "port"
        # times are the reference's own solve_time_ms (steady_clock around
        # InteriorPointSolver::solve, interior_point.cpp:883-884)
        def secs(r, wall):
            return r.solve_ms * 1e-3 if r.solve_ms > 0 else wall

        done, spent, limited = 0, 0.0, 0
        for i in idx:
            if spent >= args.cpu_seconds:
                break
            t0 = time.perf_counter()
            r = solve(lps[i], time_limit=args.cpu_lp_limit, trace_cap=0)
            spent += secs(r, time.perf_counter() - t0)
            if r.status == "OPTIMAL":
                done += 1
            else:
                limited += 1
        seen = done + limited
        out["cpu_baseline"] = {
            "value": round(done / spent, 3), "unit": "LPs/s", "cores": 1, "kind": kind,
            "sample": (f"first {seen} instances of the strided batch sample, {args.cpu_lp_limit:.0f} s limit per "
                       f"instance; {limited} stopped at the limit and count as unsolved with their time charged"),
            "host": host_info()}
        log(f"cpu lp baseline {out['cpu_baseline']}")
        # the same solver on independent instances over every usable core
        # (this repository's harness around the reference's solver)
        from concurrent.futures import ThreadPoolExecutor
        T = all_core_threads()
        sub = idx[:2 * T]
        t0 = time.perf_counter()
        with ThreadPoolExecutor(T) as ex:
            par = list(ex.map(lambda i: solve(lps[i], time_limit=args.cpu_lp_limit, trace_cap=0), sub))
        wall = time.perf_counter() - t0
        pdone = sum(r.status == "OPTIMAL" for r in par)
        out["cpu_baseline_all_cores"] = {
            "value": round(pdone / wall, 3), "unit": "LPs/s", "cores": T, "kind": kind,
            "sample": (f"first {len(sub)} instances of the strided batch sample solved concurrently by the "
                       f"reference's InteriorPointSolver, one instance per thread on {T} threads, "
                       f"{args.cpu_lp_limit:.0f} s limit each; {len(sub) - pdone} stopped at the limit and count as "
                       f"unsolved; wall time of the whole set (this repository's harness)"),
            "host": host_info()}
        log(f"cpu all-core lp baseline {out['cpu_baseline_all_cores']}")
        if kind == "reference" and not args.no_extrapolation:
            out["cpu_extrapolation"] = ref_extrapolation(O)
            log(f"cpu extrapolation {out['cpu_extrapolation']}")
        # the reference IPM on the 2,000 x 5,000 LP, bounded
        t0 = time.perf_counter()
        r = solve(lp, time_limit=args.cpu_single_limit, trace_cap=0)
        dt = secs(r, time.perf_counter() - t0)
        out["single"]["cpu_baseline"] = {
            "value": round(1.0 / dt, 4) if r.status == "OPTIMAL" else 0.0, "unit": "LPs/s", "cores": 1,
            "kind": kind, "sample": (f"one solve of the 2000x5000 LP (seed 1) under a {args.cpu_single_limit:.0f} s "
                                     f"limit: {r.status} after {r.iterations} iterations in {dt:.1f} s")}
        log(f"cpu single-LP baseline {out['single']['cpu_baseline']}")
        # the reference IPM, cold start, on the first held-out transportation LPs
        tl = []
        for sd in range(1_000_000, 1_000_002):
            t0 = time.perf_counter()
            r = solve(lpgen.transport_lp(sd, 200, 300), time_limit=args.cpu_lp_limit, trace_cap=0)
            tl.append((secs(r, time.perf_counter() - t0), r.status, r.iterations))
        out["warm_start"]["cpu_baseline"] = {
            "value": round(len(tl) / sum(t for t, _, _ in tl), 4), "unit": "LPs/s", "cores": 1, "kind": kind,
            "sample": (f"{len(tl)} held-out 200x300 transportation LPs, cold start, {args.cpu_lp_limit:.0f} s "
                       f"limit each; status/iterations " + ", ".join(f"{st}/{it}" for _, st, it in tl))}
        log(f"cpu warm-start-family baseline {out['warm_start']['cpu_baseline']}")
        # the reference IPM on the Netlib models it solves (fewer than 100 iterations)
        conv = [n for n in sorted(fx) if fx[n]["ref_ipm_iterations"] < 100 and n != "dfl001"]
        cpu, cpu_ms = {}, {}
        for n in conv:
            t0 = time.perf_counter()
            cpu[n] = solve(models[n], time_limit=args.cpu_lp_limit, trace_cap=0)
            cpu_ms[n] = 1e3 * secs(cpu[n], time.perf_counter() - t0)
        dt = sum(cpu_ms.values()) / 1e3
        gpu_t = sum(per[n]["ms"] for n in conv) / 1e3
        same = [n for n in conv if f"{cpu[n].status}/{cpu[n].iterations}" == f"{per[n]['status']}/{per[n]['iterations']}"]
        out["netlib"]["cpu_baseline"] = {
            "value": round(len(conv) / dt, 3), "unit": "LPs/s", "cores": 1, "kind": kind,
            "sample": (f"{', '.join(conv)}: {dt:.2f} s on the host ({', '.join(cpu[n].status for n in conv)}); "
                       f"the GPU solver takes {gpu_t:.3f} s on the same models"),
            "per_model_ms": {n: {"cpu": round(cpu_ms[n], 2), "gpu": per[n]["ms"],
                                 "cpu_outcome": f"{cpu[n].status}/{cpu[n].iterations}",
                                 "gpu_outcome": f"{per[n]['status']}/{per[n]['iterations']}"} for n in conv},
            "same_outcome_models": same,
            "same_outcome_cpu_s": round(sum(cpu_ms[n] for n in same) / 1e3, 4),
            "same_outcome_gpu_s": round(sum(per[n]["ms"] for n in same) / 1e3, 4)}
        log(f"cpu netlib baseline {out['netlib']['cpu_baseline']}")
        # the reference's presolve and postsolve, which Solver::solve runs around
        # the IPM (api.cpp:384-412, :176-196): their share of the end-to-end time
        # beside the GPU solve of the same models
        if O.REF is not None and hasattr(O.REF, "ref_presolve"):
            pre_ms = post_ms = 0.0
            reduced = 0
            for name in sorted(fx):
                P = O.RefPresolve(models[name])
                _, _, _, pms = P.postsolve(np.zeros(P.reduced.n))
                pre_ms += P.ms
                post_ms += pms
                reduced += (P.reduced.m, P.reduced.n) != (models[name].m, models[name].n)
            P = O.RefPresolve(lp)
            _, _, _, a_post = P.postsolve(np.zeros(P.reduced.n))
            a_gpu = single["plain_f64"]["ms"]
            gpu_netlib = 1e3 * t_all
            out["presolve"] = {
                "netlib": {"presolve_ms": round(pre_ms, 2), "postsolve_ms": round(post_ms, 2),
                           "gpu_solve_ms": round(gpu_netlib, 1), "models_reduced": reduced,
                           "share": round((pre_ms + post_ms) / (pre_ms + post_ms + gpu_netlib), 4)},
                "2000x5000": {"presolve_ms": round(P.ms, 2), "postsolve_ms": round(a_post, 3),
                              "gpu_solve_ms": a_gpu, "rows_cols_removed": [lp.m - P.reduced.m, lp.n - P.reduced.n],
                              "share": round((P.ms + a_post) / (P.ms + a_post + a_gpu), 4)},
                "kind": "reference (presolve::Presolver, AGGRESSIVE, 1 thread)"}
            log(f"presolve share {out['presolve']}")
    return out


def run_large(args, H, lpgen) -> dict:
    gold = json.load(open(os.path.join(ROOT, "tests", "golden", "highs_objectives.json")))
    out = {"generator": "powerlaw_lp(seed 3, n = 2.5 m, alpha 2.179, column counts 2..200)"}
    for m in (2000, 5000):
        lp = lpgen.powerlaw_lp(3, m, int(2.5 * m), 2.179, 2, 200)
        h = gold[f"powerlaw_{m}x{int(2.5 * m)}"]["objective"]
        H.ipm_solve(lp, pivot_rule=H.HN_PIVOT_SKIP, gap_tol=KKT_TOL)  # warm-up
        leg = {"stop": "reference test and relative gap <= 1e-8, skip pivot rule (dense normal equations)"}
        for name, prec in (("plain_f64", H.HN_PREC_F64), ("mixed", H.HN_PREC_MIXED)):
            t0 = time.perf_counter()
            r = H.ipm_solve(lp, precision=prec, pivot_rule=H.HN_PIVOT_SKIP, gap_tol=KKT_TOL)
            dt = time.perf_counter() - t0
            leg[name] = {"ms": round(1e3 * dt, 2), "status": r.status, "iterations": r.iterations,
                         "rel_err_vs_highs": abs(r.objective_value - h) / abs(h), "rel_primal": r.rel_primal,
                         "rel_dual": r.rel_dual, "rel_gap": r.rel_gap, "fp64_from_iter": r.fp64_from_iter}
        leg["mixed_speedup"] = round(leg["plain_f64"]["ms"] / leg["mixed"]["ms"], 3)
        out[f"{m}x{int(2.5 * m)}"] = leg
        log(f"large-LP generator at {m}: plain {leg['plain_f64']['ms']:.0f} ms, mixed {leg['mixed']['ms']:.0f} ms")
    if args.large_limit > 0:
        lp = lpgen.powerlaw_lp(3, 200_000, 500_000, 2.179, 2, 200)
        leg = {"nnz": lp.nnz, "time_limit_s": args.large_limit,
               "solver": "matrix-free Jacobi-PCG (dense normal equations would need 200,000^2 entries)"}
        for name, prec in (("plain_f64", H.HN_PREC_F64), ("mixed", H.HN_PREC_MIXED)):
            t0 = time.perf_counter()
            r = H.ipm_solve(lp, precision=prec, time_limit_seconds=args.large_limit, gap_tol=KKT_TOL)
            dt = time.perf_counter() - t0
            leg[name] = {"s": round(dt, 2), "status": r.status, "iterations": r.iterations, "rp_norm": r.rp_norm,
                         "rd_norm": r.rd_norm, "mu": r.mu, "rel_primal": r.rel_primal, "rel_dual": r.rel_dual,
                         "rel_gap": r.rel_gap, "cg_iterations": r.cg_iterations, "ne_solves": r.ne_solves}
            log(f"large LP 200000x500000 {name}: {r.status} after {r.iterations} iterations, {dt:.1f} s, "
                f"rp {r.rp_norm:.2e} mu {r.mu:.2e}")
        out["200000x500000"] = leg
    return out


def run_batch_bucket(args) -> dict:
    """The batched solver on the instances of the 4,096-LP workload whose m
    lies in [lo, hi): wall time, iterations, and the multiply-adds of the
    LDL' factorisations it ran (sum over instances of iterations * m^3 / 3,
    one factorisation per iteration), for counter runs per size bucket.
    With --phases the timed call also prints the library's per-phase
    workgroup time (hn_ipm_opts.verbose = 2) on stderr."""
    import hypernova_hip as H
    import lpgen
    lo, hi = (int(t) for t in args.batch_range.split(","))
    H.init(-1)
    lps = [lp for lp in (lpgen.batch_lp(i) for i in range(args.batch)) if lo <= lp.m < hi]
    H.ipm_solve_batch(lps[:8], pivot_rule=H.HN_PIVOT_SKIP, gap_tol=KKT_TOL)
    t0 = time.perf_counter()
    res = H.ipm_solve_batch(lps, pivot_rule=H.HN_PIVOT_SKIP, gap_tol=KKT_TOL, verbose=2 if args.phases else False)
    wall = time.perf_counter() - t0
    fma = sum(r.iterations * lp.m ** 3 / 3.0 for lp, r in zip(lps, res))
    return {"batch_bucket": f"m in [{lo}, {hi})", "lps": len(lps), "wall_s": round(wall, 4),
            "lps_per_s": round(len(lps) / wall, 2), "iterations_mean": round(float(np.mean([r.iterations for r in res])), 2),
            "m_mean": round(float(np.mean([lp.m for lp in lps])), 1), "factor_fma": fma,
            "factor_gflops_wall": round(2 * fma / wall / 1e9, 1),
            "kkt_1e-8": sum(max(r.rel_primal, r.rel_dual, r.rel_gap) <= KKT_TOL for r in res)}


def ref_extrapolation(O) -> dict:
    """The reference's solve time at the full sizes of the single-LP (2,000 x
    5,000) and large-LP (200,000 x 500,000) workloads, which it cannot reach
    in a bench's time: its own solve_time_ms at m = 250, 500, 1,000 of the
    same generators (n = 2.5 m), solved concurrently one per thread, then a
    least-squares fit t = a m^b in log-log space.  Labelled as an
    extrapolation everywhere; never a measurement."""
    import lpgen
    from concurrent.futures import ThreadPoolExecutor
    gens = {"single_lp_2000x5000": (lambda m: lpgen.random_lp(1, m, int(2.5 * m)), 2000),
            "large_lp_200000x500000": (lambda m: lpgen.powerlaw_lp(3, m, int(2.5 * m), 2.179, 2, 200), 200_000)}
    ms = (250, 500, 1000)
    jobs = [(name, m, make(m)) for name, (make, _) in gens.items() for m in ms]
    with ThreadPoolExecutor(len(jobs)) as ex:
        res = list(ex.map(lambda j: O.ref_ipm_solve(j[2], time_limit=120.0, trace_cap=0), jobs))
    out = {}
    for name, (_, target) in gens.items():
        pts = [(m, r) for (nm, m, _), r in zip(jobs, res) if nm == name]
        t = np.array([r.solve_ms * 1e-3 for _, r in pts])
        b, loga = np.polyfit(np.log([m for m, _ in pts]), np.log(t), 1)
        out[name] = {
            "measured": {str(m): {"s": round(r.solve_ms * 1e-3, 3), "status": r.status, "iterations": r.iterations}
                         for m, r in pts},
            "fit": f"t = {np.exp(loga):.3e} * m^{b:.2f} s",
            "extrapolated_s": float(f"{np.exp(loga) * target ** b:.3g}"),
            "at_m": target,
            "label": ("extrapolation, not a measurement: power-law fit of the reference's own solve_time_ms "
                      "(interior_point.cpp:883-884) at m = 250, 500, 1000 of the same generator, 1 thread each")}
    return out


def verify_bytes(lp, with_dual: bool) -> int:
    """Compulsory bytes of the verifier's two passes: A in CSR and A' in CSR
    (8-byte values, 4-byte indices and row pointers), the primal, bounds and
    costs, rhs and senses, the dual once per pass, one flag byte per row and
    per column."""
    m, n, z = lp.m, lp.n, lp.nnz
    rows = z * 12 + (m + 1) * 4 + n * 8 + m * 9 + m + (m * 8 if with_dual else 0)
    cols = z * 12 + (n + 1) * 4 + n * 32 + n + (m * 8 if with_dual else 0)
    return rows + cols


def run_verify(args) -> dict:
    """Post-solve verification (SolutionVerifier::verify_detailed) of a
    perturbed interior point with its dual, on the 2,000 x 5,000 LP and the
    200,000 x 500,000 power-law LP."""
    import hypernova_hip as H
    import lpgen
    H.init(-1)
    out = {}
    for name, make in (("2000x5000", lambda: lpgen.random_lp(1, 2000, 5000)),
                       ("200000x500000", lambda: lpgen.powerlaw_lp(3, 200_000, 500_000))):
        lp = make()
        g = np.random.default_rng(5)
        x = lp.x0.copy()
        idx = g.choice(lp.n, size=lp.n // 100, replace=False)
        x[idx] = -g.uniform(1e-9, 1e-3, idx.size)
        objv = float(np.dot(lp.obj, x))
        ver = H.SolutionVerifier()
        sol = H.Solution(x, lp.y0, None, objv)
        r = ver.verify_detailed(lp, sol)
        walls, devs = [], []
        for _ in range(max(3, min(args.steps, 10))):
            t0 = time.perf_counter()
            r = ver.verify_detailed(lp, sol)
            walls.append(time.perf_counter() - t0)
            devs.append(r.device_ms)
        wall_ms, dev_ms = 1e3 * float(np.median(walls)), float(np.median(devs))
        nbytes = verify_bytes(lp, True)
        out[name] = {"m": lp.m, "n": lp.n, "nnz": lp.nnz, "wall_ms": round(wall_ms, 3), "device_ms": round(dev_ms, 4),
                     "device_GBps": round(nbytes / (dev_ms * 1e-3) / 1e9, 1), "algorithmic_bytes": nbytes,
                     "result": {"feasible": r.feasible, "optimal": r.optimal, "n_violated": r.n_violated,
                                "message": r.message}}
        log(f"verify {name}: wall {wall_ms:.2f} ms, device {dev_ms:.3f} ms "
            f"({out[name]['device_GBps']} GB/s of compulsory traffic incl. the transpose build), "
            f"{r.n_violated} violated, '{r.message}'")
        if not args.no_cpu_baseline and name == "2000x5000":
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import oracle as O  # baseline only
            if O.have_reference() and hasattr(O.REF, "ref_verify"):
                rr = O.ref_verify(lp, x, lp.y0, None, objv)
                out[name]["cpu_baseline"] = {
                    "value": round(1e3 * rr.seconds, 2), "unit": "ms per verify", "cores": 1, "kind": "reference",
                    "sample": ("one SolutionVerifier::verify_detailed call on the same solution "
                               f"(same result: {rr.feasible == r.feasible and rr.optimal == r.optimal and rr.n_violated == r.n_violated})")}
                log(f"cpu verify baseline {out[name]['cpu_baseline']}")
        if name == "200000x500000":
            out[name]["cpu_baseline"] = {
                "value": None, "unit": "ms per verify", "cores": 1, "kind": "reference",
                "sample": ("not run: the reference's dual check scans all nonzeros once per column "
                           f"(n * nnz = {lp.n * lp.nnz:.2e} steps)")}
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-lp", action="store_true")
    ap.add_argument("--no-verify", action="store_true")
    ap.add_argument("--verify-only", action="store_true", help="run only the post-solve verification leg")
    ap.add_argument("--chunks-A", default="", help="fixed chunked layout of A: items,span_log2,threads")
    ap.add_argument("--chunks-At", default="", help="fixed chunked layout of A': items,span_log2,threads")
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--batch-range", default="", help="lo,hi: only the batched solver on instances with lo <= m < hi")
    ap.add_argument("--phases", action="store_true",
                    help="with --batch-range: print the per-phase workgroup time of the timed batched call")
    ap.add_argument("--seq-sample", type=int, default=64)
    ap.add_argument("--tune-per-bucket", type=int, default=64, help="sample LPs per m bucket of the batch autotune")
    ap.add_argument("--cpu-lp-limit", type=float, default=8.0)
    ap.add_argument("--ws-test", type=int, default=1000, help="held-out transportation LPs of the warm-start leg")
    ap.add_argument("--cpu-single-limit", type=float, default=20.0, help="time limit of the reference CPU solve of the 2000x5000 LP")
    ap.add_argument("--large-limit", type=float, default=15.0,
                    help="time limit (s) of each full-size 200000x500000 solve (0 skips them)")
    ap.add_argument("--no-extrapolation", action="store_true",
                    help="skip the reference solves at m = 250/500/1000 behind the labelled extrapolations")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only: a multi-GPU run (row-block sharding with "
                                   "an all-gather of x per product) is not launched", "n_gpus": args.gpus}))
        return 2
    if args.batch_range:
        print(json.dumps(run_batch_bucket(args)), flush=True)
        return 0
    if args.verify_only:
        print(json.dumps({"verify": run_verify(args)}), flush=True)
        return 0
    out = run_spmv(args)
    if not args.no_lp:
        out["lp"] = run_lp(args)
    if not args.no_verify:
        out["verify"] = run_verify(args)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
