"""The C-ABI boundary (no GPU needed): libhn_kernels.so loads, exports every
entry point declared in include/*.h, and reports errors through status codes
and hn_last_error without a device."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def declared_functions():
    names = set()
    for fn in os.listdir(INCLUDE):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(INCLUDE, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:[A-Za-z_][A-Za-z0-9_]*\s+\**)+\**\s*(hn_[A-Za-z0-9_]+)\s*\(", text, re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("hn_init", "hn_finalize", "hn_last_error", "hn_csr_upload", "hn_csr_transpose", "hn_spmv",
                 "hn_spmm", "hn_dot", "hn_axpy", "hn_norm_inf", "hn_spmv_dev"):
        assert must in names


def test_library_exports_every_declared_symbol():
    import hypernova_hip as H
    lib = C.CDLL(H.KERNEL_LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_calls_without_init_fail_with_status():
    import hypernova_hip as H
    lib = H.LIB
    out = C.c_double()
    x = (C.c_double * 4)(1, 2, 3, 4)
    if lib.hn_kernel_executions() == 0 and not _initialised(lib):
        st = lib.hn_dot(x, x, 4, C.byref(out))
        assert st == 5  # HN_ERR_NOT_INIT
        assert "hn_init" in H.last_error()


def _initialised(lib):
    f, t = C.c_size_t(), C.c_size_t()
    return lib.hn_mem_info(C.byref(f), C.byref(t)) == 0


def test_argument_validation_is_host_side():
    import hypernova_hip as H
    h = C.c_void_p()
    # row_ptr must end at nnz; checked before any device work
    rp = (C.c_uint64 * 3)(0, 1, 5)
    ci = (C.c_uint64 * 1)(0)
    v = (C.c_double * 1)(1.0)
    st = H.LIB.hn_csr_upload(v, ci, rp, 2, 2, 1, H.HN_F64, C.byref(h))
    assert st == 1 and "row_ptr" in H.last_error()
    rp2 = (C.c_uint64 * 2)(0, 1)
    ci2 = (C.c_uint64 * 1)(7)
    st = H.LIB.hn_csr_upload(v, ci2, rp2, 1, 2, 1, H.HN_F64, C.byref(h))
    assert st == 1 and "column index" in H.last_error()
    st = H.LIB.hn_csr_upload(v, ci, rp2, 1, 2, 1, 8, C.byref(h))
    assert st == 1


def test_no_gpu_here_is_reported():
    import hypernova_hip as H
    if os.environ.get("HN_EXPECT_GPU"):
        pytest.skip("GPU host")
    info = H.DeviceInfo()
    st = H.LIB.hn_init(-1, C.byref(info))
    if st == 0:
        pytest.skip("a gfx950 device is visible")
    assert st == 6 and H.last_error()


def test_noalloc_dot_rejects_oversized_n_on_the_host():
    import hypernova_hip as H
    out = C.c_double()
    x = (C.c_double * 1)(1.0)
    st = H.LIB.hn_dot_dev_noalloc(C.cast(x, C.c_void_p), C.cast(x, C.c_void_p), (1 << 62) + 1,
                                  C.cast(C.byref(out), C.c_void_p), None)
    assert st == 1 and "2^62" in H.last_error()
