"""GPU parity of the SpMV / transpose / vector kernels against the oracle.

Tolerances follow the reference's own backend tests
(tests/unit/test_gpu_backend.cpp:67-74: |got - ref| <= tol * max(1, |ref|),
tol 1e-10 on the GPU) and the FP32 bar of 1e-5 relative for the FP32 path.
Integer work (the transpose) is compared bit for bit.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

CASES = ["cpu_spmv", "cpu_spmm", "auto_spmv", "cuda_scalar", "cuda_warp", "cuda_spmm", "residency"]


def _case(g, name):
    nrows, ncols = (int(v) for v in g[f"{name}_shape"])
    return (nrows, ncols, g[f"{name}_rp"].astype(np.int64), g[f"{name}_ci"].astype(np.int64), g[f"{name}_v"],
            g[f"{name}_x"], g[f"{name}_y"])


def close(got, ref, tol):
    scale = np.maximum(1.0, np.abs(ref))
    err = np.max(np.abs(got - ref) / scale) if ref.size else 0.0
    assert err <= tol, f"max scaled error {err:.3e} > {tol:.1e}"
    return err


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("ipt", [4, 8, 16])
def test_spmv_reference_cases(hip, golden_backend, name, ipt):
    nrows, ncols, rp, ci, v, x, y_ref = _case(golden_backend, name)
    print(f"[case] spmv {name} {nrows}x{ncols} nnz={v.size} ipt={ipt}", flush=True)
    A = hip.CsrMatrix.upload(v, ci, rp, ncols, hip.HN_F64 | hip.HN_F32)
    A.set_tiling(ipt)
    close(A.spmv(x), y_ref, 1e-10)
    # FP32 path on device buffers, float accumulation
    dx = hip.DeviceArray.from_numpy(x.astype(np.float32))
    dy = hip.DeviceArray(nrows, np.float32)
    A.spmv_dev(dx, dy, hip.HN_F32)
    y32 = dy.to_numpy().astype(np.float64)
    ref32 = O.spmv(rp, ci, v.astype(np.float32).astype(np.float64), x.astype(np.float32).astype(np.float64))
    l1 = O.spmv(rp, ci, np.abs(v), np.abs(x))  # error scale of a float sum
    assert np.all(np.abs(y32 - ref32) <= 1e-5 * np.maximum(1.0, l1))


def test_spmv_empty_rows(hip):
    # test_gpu_backend.cpp:236-258: diagonal 2.0, every 4th row empty, x = 1.5
    n = 256
    rp, ci, v = [0], [], []
    for i in range(n):
        if i % 4:
            ci.append(i)
            v.append(2.0)
        rp.append(len(ci))
    x = np.full(n, 1.5)
    A = hip.CsrMatrix.upload(np.array(v), np.array(ci), np.array(rp), n)
    for ipt in (4, 8, 16):
        A.set_tiling(ipt)
        y = A.spmv(x)
        exp = np.where(np.arange(n) % 4 == 0, 0.0, 3.0)
        assert np.array_equal(y, exp)


def test_spmv_degenerate_shapes(hip):
    # empty matrix (test_gpu_backend.cpp:126-131) and all-empty rows
    A = hip.CsrMatrix.upload(np.zeros(0), np.zeros(0), np.zeros(5, np.int64), 3)
    assert np.array_equal(A.spmv(np.ones(3)), np.zeros(4))
    B = hip.HipBackend()
    B.initialize()
    assert B.spmv([], [], [0], []).size == 0


def _skewed(seed, nrows, ncols, long_rows):
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, 4, nrows)
    for r, L in long_rows:
        lens[r] = L
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    ci = np.concatenate([np.sort(rng.choice(ncols, size=L, replace=False)) for L in lens]).astype(np.int64)
    v = rng.standard_normal(rp[-1])
    x = rng.uniform(-1, 1, ncols)
    return rp, ci, v, x


@pytest.mark.parametrize("ipt", [4, 8, 16])
def test_spmv_skewed_rows_span_tiles(hip, ipt):
    # rows longer than several tiles, next to runs of empty rows
    rp, ci, v, x = _skewed(7, 5000, 20000, [(3, 9000), (4, 1), (2500, 12000), (4999, 7000)])
    print(f"[case] skewed nnz={v.size} ipt={ipt}", flush=True)
    A = hip.CsrMatrix.upload(v, ci, rp, 20000, hip.HN_F64)
    A.set_tiling(ipt)
    close(A.spmv(x), O.spmv(rp, ci, v, x), 1e-12)


@pytest.mark.parametrize("slab_log2", [4, 6, 10])
def test_spmv_slab_layout(hip, golden_backend, slab_log2):
    # the slab layout on the reference's own cases, FP64 and FP32
    for name in CASES:
        nrows, ncols, rp, ci, v, x, y_ref = _case(golden_backend, name)
        print(f"[case] slab spmv {name} {nrows}x{ncols} nnz={v.size} slab_log2={slab_log2}", flush=True)
        A = hip.CsrMatrix.upload(v, ci, rp, ncols, hip.HN_F64 | hip.HN_F32)
        A.set_slab(slab_log2)
        assert A.slab() == (slab_log2 if v.size and nrows and ncols else 0)
        close(A.spmv(x), y_ref, 1e-10)
        dx = hip.DeviceArray.from_numpy(x.astype(np.float32))
        dy = hip.DeviceArray(nrows, np.float32)
        A.spmv_dev(dx, dy, hip.HN_F32)
        ref32 = O.spmv(rp, ci, v.astype(np.float32).astype(np.float64), x.astype(np.float32).astype(np.float64))
        l1 = O.spmv(rp, ci, np.abs(v), np.abs(x))
        assert np.all(np.abs(dy.to_numpy().astype(np.float64) - ref32) <= 1e-5 * np.maximum(1.0, l1))


@pytest.mark.parametrize("slab_log2", [4, 9, 14])
def test_spmv_slab_skewed_and_blocks(hip, slab_log2):
    # rows longer than a chunk of one slab, runs of empty rows, more rows than
    # one block holds (2048), ncols not a multiple of the slab width
    rp, ci, v, x = _skewed(11, 9000, 20011, [(3, 9000), (4, 1), (2500, 12000), (2047, 3000), (2048, 5000),
                                            (8999, 7000)])
    print(f"[case] slab skewed nnz={v.size} slab_log2={slab_log2}", flush=True)
    A = hip.CsrMatrix.upload(v, ci, rp, 20011, hip.HN_F64)
    A.set_slab(slab_log2)
    ref = O.spmv(rp, ci, v, x)
    y1 = A.spmv(x)
    close(y1, ref, 1e-12)
    assert np.array_equal(A.spmv(x), y1)  # deterministic
    A.set_tiling(8)  # back to the merge path
    assert A.slab() == 0
    close(A.spmv(x), ref, 1e-12)


def test_spmv_slab_empty_rows_exact(hip):
    # test_gpu_backend.cpp:236-258 under the slab layout: exact values
    n = 5000
    rp, ci, v = [0], [], []
    for i in range(n):
        if i % 4:
            ci.append(i)
            v.append(2.0)
        rp.append(len(ci))
    A = hip.CsrMatrix.upload(np.array(v), np.array(ci), np.array(rp), n)
    A.set_slab(5)
    y = A.spmv(np.full(n, 1.5))
    assert np.array_equal(y, np.where(np.arange(n) % 4 == 0, 0.0, 3.0))


@pytest.mark.parametrize("ipt,span,nt", [(4, 4, 256), (8, 4, 1024), (4, 9, 512), (8, 19, 256), (8, 16, 1024)])
def test_spmv_chunk_layout(hip, golden_backend, ipt, span, nt):
    # the chunked stream on the reference's own cases, FP64 and FP32
    for name in CASES:
        nrows, ncols, rp, ci, v, x, y_ref = _case(golden_backend, name)
        print(f"[case] chunk spmv {name} {nrows}x{ncols} nnz={v.size} ipt={ipt} span={span} nt={nt}", flush=True)
        A = hip.CsrMatrix.upload(v, ci, rp, ncols, hip.HN_F64 | hip.HN_F32)
        A.set_chunks(ipt, span, nt)
        assert A.chunks() == ((ipt, span, nt) if v.size and nrows and ncols else (0, 0, 0))
        close(A.spmv(x), y_ref, 1e-10)
        dx = hip.DeviceArray.from_numpy(x.astype(np.float32))
        dy = hip.DeviceArray(nrows, np.float32)
        A.spmv_dev(dx, dy, hip.HN_F32)
        ref32 = O.spmv(rp, ci, v.astype(np.float32).astype(np.float64), x.astype(np.float32).astype(np.float64))
        l1 = O.spmv(rp, ci, np.abs(v), np.abs(x))
        assert np.all(np.abs(dy.to_numpy().astype(np.float64) - ref32) <= 1e-5 * np.maximum(1.0, l1))


@pytest.mark.parametrize("ipt,span,nt", [(4, 4, 256), (8, 10, 256), (4, 16, 1024), (8, 19, 512), (8, 17, 1024)])
def test_spmv_chunk_skewed_and_blocks(hip, ipt, span, nt):
    # rows longer than a chunk (9000, 12000 > 2048 nonzeros), runs of empty
    # rows, more rows than one block holds (2048), ncols not a power of two;
    # narrow spans force chunks cut by the span window
    rp, ci, v, x = _skewed(11, 9000, 20011, [(3, 9000), (4, 1), (2500, 12000), (2047, 3000), (2048, 5000),
                                            (8999, 7000)])
    print(f"[case] chunk skewed nnz={v.size} ipt={ipt} span={span} nt={nt}", flush=True)
    A = hip.CsrMatrix.upload(v, ci, rp, 20011, hip.HN_F64 | hip.HN_F32)
    A.set_chunks(ipt, span, nt)
    ref = O.spmv(rp, ci, v, x)
    y1 = A.spmv(x)
    close(y1, ref, 1e-12)
    assert np.array_equal(A.spmv(x), y1)  # deterministic
    dx = hip.DeviceArray.from_numpy(x.astype(np.float32))
    dy = hip.DeviceArray(9000, np.float32)
    A.spmv_dev(dx, dy, hip.HN_F32)
    ref32 = O.spmv(rp, ci, v.astype(np.float32).astype(np.float64), x.astype(np.float32).astype(np.float64))
    l1 = O.spmv(rp, ci, np.abs(v), np.abs(x))
    assert np.all(np.abs(dy.to_numpy().astype(np.float64) - ref32) <= 1e-5 * np.maximum(1.0, l1))
    A.set_tiling(8)  # back to the merge path
    assert A.chunks() == (0, 0, 0) and A.slab() == 0
    close(A.spmv(x), ref, 1e-12)


def test_spmv_chunk_single_long_row_and_exact(hip):
    # one row holding every nonzero (runs across all threads, waves and
    # chunks), then the empty-row diagonal case of test_gpu_backend.cpp:236-258
    n = 50000
    rng = np.random.default_rng(5)
    ci = np.sort(rng.choice(n, 20000, replace=False))
    v = rng.standard_normal(20000)
    rp = np.array([0, 20000, 20000])
    x = rng.uniform(-1, 1, n)
    A = hip.CsrMatrix.upload(v, ci, rp, n)
    for ipt, span, nt in ((4, 12, 256), (8, 19, 1024)):
        A.set_chunks(ipt, span, nt)
        close(A.spmv(x), O.spmv(rp, ci, v, x), 1e-12)
    n = 5000
    rp, ci, v = [0], [], []
    for i in range(n):
        if i % 4:
            ci.append(i)
            v.append(2.0)
        rp.append(len(ci))
    A = hip.CsrMatrix.upload(np.array(v), np.array(ci), np.array(rp), n)
    A.set_chunks(8, 5, 512)
    y = A.spmv(np.full(n, 1.5))
    assert np.array_equal(y, np.where(np.arange(n) % 4 == 0, 0.0, 3.0))


def test_transpose_bit_exact(hip, golden_backend):
    for name in CASES:
        nrows, ncols, rp, ci, v, x, _ = _case(golden_backend, name)
        A = hip.CsrMatrix.upload(v, ci, rp, ncols, hip.HN_F64 | hip.HN_F32)
        At = A.transpose()
        tv, tc, tr = At.download()
        orp, oci, ov = O.csr_transpose(rp, ci, v, ncols)
        assert np.array_equal(tr.astype(np.int64), orp)
        assert np.array_equal(tc.astype(np.int64), oci)
        assert np.array_equal(tv, ov)
        # A' y through the transpose == scatter matvec_t within 1e-12
        w = np.linspace(-1, 1, nrows)
        close(At.spmv(w), O.spmv_t(rp, ci, v, w, ncols), 1e-12)


def test_transpose_long_columns(hip):
    # columns longer than 16 / 4096 elements exercise the LDS and merge sorts
    rng = np.random.default_rng(11)
    nrows, ncols = 20000, 50
    rows = []
    for i in range(nrows):
        k = rng.integers(1, 4)
        rows.append(np.sort(rng.choice(ncols, size=k, replace=False)))
    rp = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int64)
    ci = np.concatenate(rows).astype(np.int64)
    v = rng.standard_normal(ci.size)
    A = hip.CsrMatrix.upload(v, ci, rp, ncols)
    tv, tc, tr = A.transpose().download()
    orp, oci, ov = O.csr_transpose(rp, ci, v, ncols)
    assert np.array_equal(tr.astype(np.int64), orp)
    assert np.array_equal(tc.astype(np.int64), oci)
    assert np.array_equal(tv, ov)


def test_double_transpose_identity(hip):
    rp, ci, v, x = _skewed(3, 3000, 7000, [(10, 5000)])
    A = hip.CsrMatrix.upload(v, ci, rp, 7000)
    v2, c2, r2 = A.transpose().transpose().download()
    assert np.array_equal(r2.astype(np.int64), rp) and np.array_equal(c2.astype(np.int64), ci)
    assert np.array_equal(v2, v)


def test_spmm(hip, golden_backend):
    for name in ("cpu_spmm", "cuda_spmm"):
        nrows, ncols, rp, ci, v, x, _ = _case(golden_backend, name)
        B = golden_backend[f"{name}_B"]
        k = int(golden_backend[f"{name}_k"][0])
        A = hip.CsrMatrix.upload(v, ci, rp, ncols)
        close(A.spmm(B, k), O.spmm(rp, ci, v, B, k), 1e-10)


def test_vector_ops(hip):
    rng = np.random.default_rng(5)
    be = hip.HipBackend()
    be.initialize()
    for n in (1, 63, 10000, 1_000_003):
        a = rng.standard_normal(n)
        b = rng.standard_normal(n)
        d = be.dot(a, b)
        assert abs(d - O.dot(a, b)) <= 1e-12 * max(1.0, float(np.abs(a * b).sum()))
        a[n // 2] = np.nan
        assert be.norm_inf(a) == O.norm_inf(a)  # order-free, NaN ignored: bit-exact
        y = rng.standard_normal(n)
        y2 = y.copy()
        be.axpy(0.37, b, y)
        O.OR.or_axpy(n, 0.37, b.ctypes.data_as(O._dp), y2.ctypes.data_as(O._dp))
        assert np.array_equal(y, y2)
    assert be.was_gpu_kernel_executed()


def test_backend_residency_and_mirror(hip, golden_backend):
    nrows, ncols, rp, ci, v, x, y_ref = _case(golden_backend, "residency")
    be = hip.HipBackend()
    be.initialize()
    assert be.type() == hip.ComputeBackendType.HIP and be.is_available()
    vals, cols, rows = list(v), list(ci), list(rp)
    y1 = be.spmv(vals, cols, rows, x)
    free1 = be.free_memory()
    y2 = be.spmv(vals, cols, rows, x)
    assert be.free_memory() <= free1 + (1 << 20)
    close(y1, y_ref, 1e-10)
    assert np.array_equal(y1, y2)
    vals[0] += 1.0  # same containers, new contents: must not return stale results
    y3 = be.spmv(vals, cols, rows, x)
    assert abs((y3 - y1)[0] - x[ci[0]]) < 1e-12
    be.finalize()


def test_autotune_keeps_results(hip):
    rp, ci, v, x = _skewed(9, 20000, 30000, [(100, 6000)])
    A = hip.CsrMatrix.upload(v, ci, rp, 30000, hip.HN_F64 | hip.HN_F32)
    us = A.autotune(5)
    assert us > 0 and A.tiling() in (4, 8, 16) and A.slab() in (0, 14, 15, 16, 17, 18, 19)
    ci_, sp_, nt_ = A.chunks()
    assert (ci_, sp_, nt_) == (0, 0, 0) or (ci_ in (4, 8) and 17 <= sp_ <= 19 and nt_ in (256, 512, 1024))
    assert not (A.slab() and A.chunks()[0])
    close(A.spmv(x), O.spmv(rp, ci, v, x), 1e-12)
