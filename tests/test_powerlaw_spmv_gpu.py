"""The SpMV workload's own generator on the GPU: the Zipf(1.2) power-law
matrix (row lengths capped at 10,000, rescaled to the target nonzero count,
uniform distinct columns), through every layout the autotune chooses from,
against the oracle's FP64 products.

Bars: FP64 within 1e-12 of the CPU FP64 product, scaled by max(1, |y|);
FP32 (float values, float x, float accumulation) within 1e-5 of the same
product in FP64 arithmetic over the float-rounded inputs, scaled by the
l1 row sum; the transpose bit-exact."""
import numpy as np
import pytest

import lpgen
import oracle as O

pytestmark = pytest.mark.gpu

LAYOUTS = [("merge", 4), ("merge", 8), ("merge", 16), ("slab", 14), ("slab", 17),
           ("chunk", (4, 17, 1024)), ("chunk", (8, 18, 256)), ("chunk", (4, 19, 512)), ("chunk", (4, 19, 1024))]


@pytest.fixture(scope="module")
def reduced():
    # the same rule at a tenth of the size: 200,000 x 500,000, 4M nonzeros
    c = lpgen.powerlaw_spmv(2, 200_000, 500_000, 4_000_000)
    rp = c.row_ptr
    ci = c.col.astype(np.int64)
    return c, rp, ci, O.spmv(rp, ci, c.val, c.x), O.spmv_t(rp, ci, c.val, c.w, c.ncols)


def _set(A, kind, arg):
    if kind == "merge":
        A.set_tiling(arg)
    elif kind == "slab":
        A.set_slab(arg)
    else:
        A.set_chunks(*arg)


def _err(got, ref):
    return float(np.max(np.abs(got - ref) / np.maximum(1.0, np.abs(ref))))


def _f32_ok(hip, A, c, rp, ci, nrows, x):
    dx = hip.DeviceArray.from_numpy(x.astype(np.float32))
    dy = hip.DeviceArray(nrows, np.float32)
    A.spmv_dev(dx, dy, hip.HN_F32)
    ref32 = O.spmv(rp, ci, c.astype(np.float32).astype(np.float64), x.astype(np.float32).astype(np.float64))
    l1 = O.spmv(rp, ci, np.abs(c), np.abs(x))
    return float(np.max(np.abs(dy.to_numpy().astype(np.float64) - ref32) / np.maximum(1.0, l1)))


@pytest.mark.parametrize("kind,arg", LAYOUTS, ids=[f"{k}-{a}" for k, a in LAYOUTS])
def test_powerlaw_reduced_every_layout(hip, reduced, kind, arg):
    c, rp, ci, yref, zref = reduced
    print(f"[case] powerlaw 200kx500k nnz={c.nnz} layout {kind} {arg}", flush=True)
    A = hip.CsrMatrix.upload(c.val, ci, rp, c.ncols, hip.HN_F64 | hip.HN_F32)
    At = A.transpose()
    _set(A, kind, arg)
    _set(At, kind, arg)
    e1, e2 = _err(A.spmv(c.x), yref), _err(At.spmv(c.w), zref)
    f1 = _f32_ok(hip, A, c.val, rp, ci, c.nrows, c.x)
    print(f"  A x {e1:.2e}  A' w {e2:.2e}  FP32 A x {f1:.2e}", flush=True)
    assert e1 <= 1e-12 and e2 <= 1e-12 and f1 <= 1e-5


def test_powerlaw_reduced_autotuned_and_transpose(hip, reduced):
    c, rp, ci, yref, zref = reduced
    A = hip.CsrMatrix.upload(c.val, ci, rp, c.ncols, hip.HN_F64 | hip.HN_F32)
    At = A.transpose()
    tv, tc, tr = At.download()
    orp, oci, ov = O.csr_transpose(rp, ci, c.val, c.ncols)
    assert np.array_equal(tr.astype(np.int64), orp) and np.array_equal(tc.astype(np.int64), oci)
    assert np.array_equal(tv, ov)
    A.autotune(3)
    At.autotune(3)
    print(f"[case] autotuned: A tiling {A.tiling()} slab {A.slab()} chunks {A.chunks()}; "
          f"A' tiling {At.tiling()} slab {At.slab()} chunks {At.chunks()}", flush=True)
    assert _err(A.spmv(c.x), yref) <= 1e-12 and _err(At.spmv(c.w), zref) <= 1e-12


def test_powerlaw_full_size_benchmarked_matrix(hip):
    # the benchmark's matrix itself: 2,000,000 x 5,000,000, 40M nonzeros, seed 2,
    # under the layouts the benchmark times
    c = lpgen.powerlaw_spmv(2)
    rp, ci = c.row_ptr, c.col.astype(np.int64)
    print(f"[case] powerlaw full size nnz={c.nnz}", flush=True)
    assert c.nnz == 40_000_000 and int(np.diff(rp).max()) <= 10_000
    A = hip.CsrMatrix.upload(c.val, ci, rp, c.ncols, hip.HN_F64 | hip.HN_F32)
    At = A.transpose()
    A.set_chunks(4, 19, 1024)
    e1 = _err(A.spmv(c.x), O.spmv(rp, ci, c.val, c.x))
    f1 = _f32_ok(hip, A, c.val, rp, ci, c.nrows, c.x)
    print(f"  A x {e1:.2e}  FP32 A x {f1:.2e}", flush=True)
    assert e1 <= 1e-12 and f1 <= 1e-5
    # The transpose has 5,000,000 short rows: for every block to be resident
    # at once its blocks hold more than 8 rows per thread (about 10,000 rows
    # at 512 threads, 20,000 at 1024), so the packed word gives the row 14 or
    # 15 bits and the chunk span shrinks to 2^18 or 2^17 columns.
    zref = O.spmv_t(rp, ci, c.val, c.w, c.ncols)
    trp, trow, tval = _t_pattern(rp, ci, c.val, c.ncols)
    narrowed = 0
    for layout in ((4, 19, 512), (4, 19, 1024), (8, 19, 256)):
        At.set_chunks(*layout)
        got = At.chunks()
        e2 = _err(At.spmv(c.w), zref)
        g2 = _f32_ok(hip, At, tval, trp, trow, c.ncols, c.w)
        print(f"  A' w {layout} -> {got}: {e2:.2e}  FP32 {g2:.2e}", flush=True)
        assert got[0] == layout[0] and got[2] == layout[2] and 17 <= got[1] <= layout[1]
        narrowed += got[1] < layout[1]
        assert e2 <= 1e-12 and g2 <= 1e-5
    assert narrowed >= 2  # the tall blocks were exercised


def _t_pattern(rp, ci, val, ncols):
    # CSR of the transpose (row pointer, column indices, values), each row
    # in the original row order
    order = np.argsort(ci, kind="stable")
    trp = np.concatenate(([0], np.cumsum(np.bincount(ci, minlength=ncols)))).astype(np.int64)
    rows = np.repeat(np.arange(len(rp) - 1, dtype=np.int64), np.diff(rp))
    return trp, rows[order], val[order]
