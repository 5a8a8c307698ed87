"""Host-forced failures of the C-ABI (hn_inject_failures).

The reference's backend retries a failed device call on the CPU for that
call only (AutoBackend::spmv, core/execution/gpu_backend.cpp:688-694); its
tests reach that path by disabling the GPU.  Here the failure is forced from
the host with a device present, so the caller's fallback runs for real:
  * CPU: the injected code and message come back from an entry point before
    anything touches a device, the count is consumed call by call, and 0
    clears it;
  * GPU: a failed SpMV is retried on the CPU (the oracle's product) and the
    next call runs on the device again; a failed IPM solve leaves the
    library usable; the C++ plugin's own fallback case runs in
    tests/cpp/test_hip_backend.cpp (InjectedFailureFallsBackToCpu)."""
import ctypes as C

import numpy as np
import pytest

import hypernova_hip as H
import lpgen
import oracle as O


def _dot(n=3):
    a = np.arange(1.0, n + 1.0)
    out = C.c_double()
    st = H.LIB.hn_dot(a.ctypes.data_as(H._dp), a.ctypes.data_as(H._dp), n, C.byref(out))
    return st, out.value


def test_injected_failures_are_consumed_in_order():
    H.inject_failures(2, 2)
    for _ in range(2):
        st, _ = _dot()
        assert st == 2 and "injected failure" in H.last_error()
    H.inject_failures(1, 4)
    assert _dot()[0] == 4
    H.inject_failures(3, 2)
    H.inject_failures(0, 2)  # cleared
    st, _ = _dot()
    assert st != 2 or "injected failure" not in H.last_error()


def test_injected_code_must_be_an_error():
    with pytest.raises(H.HnError):
        H.inject_failures(1, 0)


@pytest.mark.gpu
def test_spmv_falls_back_to_cpu_for_the_failed_call(hip):
    case = lpgen.powerlaw_spmv(5, 3000, 7000, 40000, cap=500)
    A = hip.CsrMatrix.upload(case.val, case.col, case.row_ptr, case.ncols, hip.HN_F64)
    ref = O.spmv(case.row_ptr, case.col.astype(np.int64), case.val, case.x)

    def auto_spmv():  # AutoBackend::spmv's rule: a failed device call runs on the CPU
        try:
            return A.spmv(case.x), "device"
        except hip.HnError:
            return O.spmv(case.row_ptr, case.col.astype(np.int64), case.val, case.x), "cpu"

    hip.inject_failures(1)
    y1, w1 = auto_spmv()
    y2, w2 = auto_spmv()
    assert (w1, w2) == ("cpu", "device")
    assert np.array_equal(y1, ref)
    assert np.max(np.abs(y2 - ref) / np.maximum(1.0, np.abs(ref))) <= 1e-12


@pytest.mark.gpu
def test_ipm_solve_fails_cleanly_and_the_library_stays_usable(hip):
    lp = lpgen.random_lp(2, 60, 150)
    hip.inject_failures(1)
    with pytest.raises(hip.HnError, match="injected failure"):
        hip.ipm_solve(lp)
    r = hip.ipm_solve(lp)
    ref = O.ipm_solve(lp)
    assert r.status == ref.status and r.iterations == ref.iterations
