"""Python host side of the MI355X LP hot path (ctypes over ``libhn_kernels.so``).

The classes mirror the reference's plugin interface so that code written
against ``hypernova::execution::IComputeBackend`` (``core/execution/gpu_backend.hpp:18-51``)
reads the same here:

* :class:`HipBackend` - ``type/name/is_available/initialize/finalize/spmv/spmm/
  dot/axpy/norm_inf/device_memory/free_memory/was_gpu_kernel_executed/
  gpu_kernel_executions/reset_execution_stats``.  Like ``CudaDriverBackend``
  (``gpu_backend.cpp:12-14``) every failure after initialisation raises
  ``RuntimeError``; nothing falls back to the CPU.
* :class:`CsrMatrix` - an explicit device-resident matrix handle (the
  reference keys residency by the host ``values.data()`` pointer,
  ``gpu_backend.cpp:287-306``; an explicit handle cannot go stale).
* :class:`DeviceArray` - a device buffer for the ``_dev`` entry points.

The library is loaded from ``lib/`` next to this file; if it is missing the
import raises ``ImportError`` (build it with ``python build.py``).
"""
from __future__ import annotations

import ctypes as C
import enum
from dataclasses import dataclass
import os
from typing import Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, "lib")
KERNEL_LIB = os.path.join(LIBDIR, "libhn_kernels.so")
GEN_LIB = os.path.join(LIBDIR, "libhn_gen.so")

HN_OK = 0
HN_F64 = 1
HN_F32 = 2

_u64 = C.c_uint64
_i64 = C.c_int64
_dp = C.POINTER(C.c_double)
_u64p = C.POINTER(C.c_uint64)
_vp = C.c_void_p


class DeviceInfo(C.Structure):
    _fields_ = [("name", C.c_char * 256), ("arch", C.c_char * 64), ("total_memory", C.c_uint64),
                ("compute_units", C.c_int), ("device", C.c_int)]


def _load_kernels() -> C.CDLL:
    if not os.path.exists(KERNEL_LIB):
        raise ImportError(f"{KERNEL_LIB} is not built; run build.py (there is no CPU fallback)")
    lib = C.CDLL(KERNEL_LIB)
    sig = {
        "hn_init": (C.c_int, [C.c_int, C.POINTER(DeviceInfo)]),
        "hn_finalize": (C.c_int, []),
        "hn_last_error": (C.c_int, [C.c_char_p, C.c_size_t]),
        "hn_mem_info": (C.c_int, [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
        "hn_kernel_executions": (C.c_uint64, []),
        "hn_reset_execution_stats": (None, []),
        "hn_csr_upload": (C.c_int, [_dp, _u64p, _u64p, _u64, _u64, _u64, C.c_int, C.POINTER(_vp)]),
        "hn_csr_transpose": (C.c_int, [_vp, C.POINTER(_vp)]),
        "hn_csr_adat_pattern": (C.c_int, [_vp, C.POINTER(_vp)]),
        "hn_inject_failures": (C.c_int, [C.c_uint32, C.c_int]),
        "hn_sldl_analyze": (C.c_int, [_u64, _u64p, _u64p, C.c_double, C.POINTER(_vp)]),
        "hn_sldl_sizes": (C.c_int, [_vp, _u64p, _u64p, _u64p, C.POINTER(C.c_int)]),
        "hn_sldl_symbolic": (C.c_int, [_vp, _u64p, _u64p, _u64p, _u64p, _u64p]),
        "hn_sldl_factor": (C.c_int, [_vp, _vp, C.c_double, C.POINTER(C.c_int)]),
        "hn_sldl_factors": (C.c_int, [_vp, _dp, _dp]),
        "hn_sldl_solve_dev": (C.c_int, [_vp, _vp, _vp, _vp]),
        "hn_sldl_free": (C.c_int, [_vp]),
        "hn_csr_free": (C.c_int, [_vp]),
        "hn_csr_dims": (C.c_int, [_vp, _u64p, _u64p, _u64p]),
        "hn_csr_download": (C.c_int, [_vp, _dp, _u64p, _u64p]),
        "hn_csr_set_tiling": (C.c_int, [_vp, C.c_int]),
        "hn_csr_get_tiling": (C.c_int, [_vp, C.POINTER(C.c_int)]),
        "hn_csr_set_slab": (C.c_int, [_vp, C.c_int]),
        "hn_csr_get_slab": (C.c_int, [_vp, C.POINTER(C.c_int)]),
        "hn_csr_set_chunks": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int]),
        "hn_csr_get_chunks": (C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "hn_csr_autotune": (C.c_int, [_vp, C.c_int, _dp]),
        "hn_spmv": (C.c_int, [_vp, _dp, _dp]),
        "hn_spmm": (C.c_int, [_vp, _dp, _dp, _u64]),
        "hn_dot": (C.c_int, [_dp, _dp, _u64, _dp]),
        "hn_axpy": (C.c_int, [C.c_double, _dp, _dp, _u64]),
        "hn_norm_inf": (C.c_int, [_dp, _u64, _dp]),
        "hn_spmv_dev": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp]),
        "hn_spmm_dev": (C.c_int, [_vp, _vp, _vp, _u64, _vp]),
        "hn_dot_dev": (C.c_int, [_vp, _vp, _u64, _vp, _vp]),
        "hn_dot_dev_noalloc": (C.c_int, [_vp, _vp, _u64, _vp, _vp]),
        "hn_norm_inf_dev": (C.c_int, [_vp, _u64, _vp, _vp]),
        "hn_ratio_test_dev": (C.c_int, [_vp, _vp, _vp, _vp, _u64, _vp, _vp]),
        "hn_axpy_dev": (C.c_int, [C.c_double, _vp, _vp, _u64, _vp]),
        "hn_dev_alloc": (C.c_int, [C.c_size_t, C.POINTER(_vp)]),
        "hn_dev_free": (C.c_int, [_vp]),
        "hn_memcpy_h2d": (C.c_int, [_vp, _vp, C.c_size_t]),
        "hn_memcpy_d2h": (C.c_int, [_vp, _vp, C.c_size_t]),
        "hn_stream_sync": (C.c_int, [_vp]),
        "hn_copy_dev": (C.c_int, [_vp, _vp, C.c_size_t, _vp]),
        "hn_event_create": (C.c_int, [C.POINTER(_vp)]),
        "hn_event_destroy": (C.c_int, [_vp]),
        "hn_event_record": (C.c_int, [_vp, _vp]),
        "hn_event_elapsed_ms": (C.c_int, [_vp, _vp, C.POINTER(C.c_float)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


LIB = _load_kernels()


class HnError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[hn status {code}] {msg}")
        self.code = code


def last_error() -> str:
    buf = C.create_string_buffer(1024)
    LIB.hn_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def check(status: int) -> None:
    if status != HN_OK:
        raise HnError(status, last_error())


def inject_failures(count: int, code: int = 2) -> None:
    """hn_inject_failures: the next `count` library calls that use the device
    fail with `code` (default HN_ERR_HIP) before touching it; 0 clears.  For
    exercising a caller's CPU fallback with a device present."""
    check(LIB.hn_inject_failures(count, code))


def _dptr(a: np.ndarray) -> _dp:
    return a.ctypes.data_as(_dp)


def _u64ptr(a: np.ndarray) -> _u64p:
    return a.ctypes.data_as(_u64p)


_device_info: Optional[DeviceInfo] = None


def init(device: int = -1) -> DeviceInfo:
    """Selects the device and creates the library stream (idempotent)."""
    global _device_info
    info = DeviceInfo()
    check(LIB.hn_init(device, C.byref(info)))
    _device_info = info
    return info


def finalize() -> None:
    check(LIB.hn_finalize())


def mem_info() -> tuple[int, int]:
    f, t = C.c_size_t(), C.c_size_t()
    check(LIB.hn_mem_info(C.byref(f), C.byref(t)))
    return f.value, t.value


class DeviceArray:
    """A device buffer holding a 1-D numpy dtype (float64, float32, int32...)."""

    def __init__(self, n: int, dtype=np.float64):
        self.dtype = np.dtype(dtype)
        self.n = int(n)
        self.nbytes = max(self.n * self.dtype.itemsize, 1)
        p = _vp()
        check(LIB.hn_dev_alloc(self.nbytes, C.byref(p)))
        self.ptr = p

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        if a.size:
            check(LIB.hn_memcpy_h2d(d.ptr, a.ctypes.data_as(_vp), a.nbytes))
        return d

    def to_numpy(self) -> np.ndarray:
        out = np.empty(self.n, dtype=self.dtype)
        if self.n:
            check(LIB.hn_memcpy_d2h(out.ctypes.data_as(_vp), self.ptr, out.nbytes))
        return out

    def copy_from(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        if a.size:
            check(LIB.hn_memcpy_h2d(self.ptr, a.ctypes.data_as(_vp), a.nbytes))

    def free(self) -> None:
        if self.ptr:
            LIB.hn_dev_free(self.ptr)
            self.ptr = _vp()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class SparseLDL:
    """SparseCholesky (core/numerical/factorization.hpp:162-205): the exact
    minimum-degree analysis runs on the host at construction
    (hn_sldl_analyze, no device needed), the numeric LDL' and the solves on
    the device (hn_sldl_factor, hn_sldl_solve_dev)."""

    def __init__(self, row_ptr, col, work_limit: float = 0.0):
        rp = np.ascontiguousarray(row_ptr, np.uint64)
        ci = np.ascontiguousarray(col, np.uint64)
        self.h = _vp()
        check(LIB.hn_sldl_analyze(rp.size - 1, rp.ctypes.data_as(_u64p), ci.ctypes.data_as(_u64p),
                                  float(work_limit), C.byref(self.h)))

    def sizes(self) -> tuple[int, int, int, bool]:
        """(n, entries of L, dependency levels, aborted)"""
        n, nnz, lv = _u64(), _u64(), _u64()
        ab = C.c_int()
        check(LIB.hn_sldl_sizes(self.h, C.byref(n), C.byref(nnz), C.byref(lv), C.byref(ab)))
        return int(n.value), int(nnz.value), int(lv.value), bool(ab.value)

    def symbolic(self) -> dict:
        n, nnz, _, _ = self.sizes()
        out = {k: np.zeros(n, np.uint64) for k in ("perm", "parent", "postorder")}
        out["l_row_ptr"] = np.zeros(n + 1, np.uint64)
        out["l_col"] = np.zeros(max(nnz, 1), np.uint64)
        check(LIB.hn_sldl_symbolic(self.h, *(out[k].ctypes.data_as(_u64p)
                                             for k in ("perm", "l_row_ptr", "l_col", "parent", "postorder"))))
        out["l_col"] = out["l_col"][:nnz]
        return {k: v.astype(np.int64) for k, v in out.items()}

    def factor(self, a: "CsrMatrix", singular_tol: float = 1e-10) -> int:
        """Numeric LDL' of the device matrix a; returns the number of clamped pivots."""
        c = C.c_int()
        check(LIB.hn_sldl_factor(self.h, a.h, singular_tol, C.byref(c)))
        return int(c.value)

    def factors(self) -> tuple[np.ndarray, np.ndarray]:
        n, nnz, _, _ = self.sizes()
        L, D = np.zeros(max(nnz, 1)), np.zeros(max(n, 1))
        check(LIB.hn_sldl_factors(self.h, _dptr(L), _dptr(D)))
        return L[:nnz], D[:n]

    def solve(self, b) -> np.ndarray:
        db = DeviceArray.from_numpy(np.ascontiguousarray(b, np.float64))
        dx = DeviceArray(db.n)
        check(LIB.hn_sldl_solve_dev(self.h, db.ptr, dx.ptr, None))
        return dx.to_numpy()

    def __del__(self):
        if getattr(self, "h", None) and LIB is not None:
            LIB.hn_sldl_free(self.h)
            self.h = None


class CsrMatrix:
    """Device-resident CSR matrix (explicit handle)."""

    def __init__(self, handle: _vp, owner: bool = True):
        self.h = handle
        self._owner = owner
        r, c, z = C.c_uint64(), C.c_uint64(), C.c_uint64()
        check(LIB.hn_csr_dims(self.h, C.byref(r), C.byref(c), C.byref(z)))
        self.nrows, self.ncols, self.nnz = r.value, c.value, z.value

    @classmethod
    def upload(cls, values: np.ndarray, col_indices: np.ndarray, row_ptr: np.ndarray,
               ncols: int, flags: int = HN_F64) -> "CsrMatrix":
        values = np.ascontiguousarray(values, dtype=np.float64)
        col = np.ascontiguousarray(col_indices, dtype=np.uint64)
        rp = np.ascontiguousarray(row_ptr, dtype=np.uint64)
        h = _vp()
        check(LIB.hn_csr_upload(_dptr(values), _u64ptr(col), _u64ptr(rp), rp.size - 1, ncols,
                                values.size, flags, C.byref(h)))
        return cls(h)

    def transpose(self) -> "CsrMatrix":
        h = _vp()
        check(LIB.hn_csr_transpose(self.h, C.byref(h)))
        return CsrMatrix(h)

    def adat_pattern(self) -> "CsrMatrix":
        """Pattern of A D A' (D > 0) built on the device (hn_csr_adat_pattern)."""
        h = _vp()
        check(LIB.hn_csr_adat_pattern(self.h, C.byref(h)))
        return CsrMatrix(h)

    def download(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        v = np.empty(self.nnz, np.float64)
        c = np.empty(self.nnz, np.uint64)
        r = np.empty(self.nrows + 1, np.uint64)
        check(LIB.hn_csr_download(self.h, _dptr(v), _u64ptr(c), _u64ptr(r)))
        return v, c, r

    def set_tiling(self, items_per_thread: int) -> None:
        check(LIB.hn_csr_set_tiling(self.h, items_per_thread))

    def tiling(self) -> int:
        v = C.c_int()
        check(LIB.hn_csr_get_tiling(self.h, C.byref(v)))
        return v.value

    def set_slab(self, slab_log2: int) -> None:
        """Slab layout with 2**slab_log2 columns per slab (0: merge path)."""
        check(LIB.hn_csr_set_slab(self.h, slab_log2))

    def slab(self) -> int:
        v = C.c_int()
        check(LIB.hn_csr_get_slab(self.h, C.byref(v)))
        return v.value

    def set_chunks(self, items_per_thread: int, span_log2: int = 19, threads: int = 256) -> None:
        """Chunked stream: workgroups of `threads` threads, threads *
        items_per_thread nonzeros per chunk spanning at most 2**span_log2
        columns (items_per_thread 0: drop it)."""
        check(LIB.hn_csr_set_chunks(self.h, items_per_thread, span_log2, threads))

    def chunks(self) -> tuple[int, int, int]:
        """(items_per_thread, span_log2, threads) of the chunked stream,
        (0, 0, 0) if off."""
        v, sp, nt = C.c_int(), C.c_int(), C.c_int()
        check(LIB.hn_csr_get_chunks(self.h, C.byref(v), C.byref(sp), C.byref(nt)))
        return v.value, sp.value, nt.value

    def autotune(self, iters: int = 20) -> float:
        us = C.c_double()
        check(LIB.hn_csr_autotune(self.h, iters, C.byref(us)))
        return us.value

    def spmv(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float64)
        if x.size != self.ncols:
            raise ValueError("x has the wrong length")
        y = np.empty(self.nrows, np.float64)
        check(LIB.hn_spmv(self.h, _dptr(x), _dptr(y)))
        return y

    def spmm(self, B: np.ndarray, k: int) -> np.ndarray:
        B = np.ascontiguousarray(B, dtype=np.float64).ravel()
        if B.size != self.ncols * k:
            raise ValueError("B has the wrong size")
        Cm = np.empty(self.nrows * k, np.float64)
        check(LIB.hn_spmm(self.h, _dptr(B), _dptr(Cm), k))
        return Cm

    def spmv_dev(self, x: DeviceArray, y: DeviceArray, prec: int = HN_F64, stream=None) -> None:
        check(LIB.hn_spmv_dev(self.h, prec, x.ptr, y.ptr, stream))

    def free(self) -> None:
        if self._owner and self.h:
            LIB.hn_csr_free(self.h)
            self.h = _vp()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Event:
    def __init__(self):
        p = _vp()
        check(LIB.hn_event_create(C.byref(p)))
        self.p = p

    def record(self, stream=None) -> None:
        check(LIB.hn_event_record(self.p, stream))

    def elapsed_ms(self, stop: "Event") -> float:
        ms = C.c_float()
        check(LIB.hn_event_elapsed_ms(self.p, stop.p, C.byref(ms)))
        return float(ms.value)

    def __del__(self):
        try:
            LIB.hn_event_destroy(self.p)
        except Exception:
            pass


def stream_sync(stream=None) -> None:
    check(LIB.hn_stream_sync(stream))


# ---------------------------------------------------------------------------
# IComputeBackend mirror
# ---------------------------------------------------------------------------
class ComputeBackendType(enum.IntEnum):
    """Same enumerators, same order as gpu_backend.hpp:10-16."""
    CPU = 0
    CUDA = 1
    HIP = 2
    SYCL = 3
    AUTO = 4


class HipBackend:
    """``IComputeBackend`` on MI355X (the slot ``ComputeBackendType::HIP``
    that the reference declares but maps to CPU, gpu_backend.cpp:926-933).

    ``spmv(values, col_indices, row_ptr, x, y)`` keeps the matrix resident.
    The key is the identity of the three arrays plus their shape, and a
    change of any element's bytes is caught by comparing a checksum of the
    host arrays, so a different matrix at the same addresses is re-uploaded
    (the reference's pointer-only key would return stale results).
    """

    def __init__(self):
        self._available = False
        self._resident: dict = {}
        self._name = "HIP"

    def type(self) -> ComputeBackendType:
        return ComputeBackendType.HIP

    def name(self) -> str:
        return self._name

    def is_available(self) -> bool:
        return self._available

    def initialize(self) -> None:
        info = init(-1)
        self._available = True
        self._name = f"HIP({info.name.decode(errors='replace')})"

    def finalize(self) -> None:
        for m in self._resident.values():
            m[1].free()
        self._resident.clear()

    def _require(self, op: str) -> None:
        if not self._available:
            raise RuntimeError(f"HIP backend not available for {op}: {last_error()}")

    def _matrix(self, values, col_indices, row_ptr, ncols: int) -> CsrMatrix:
        v = np.ascontiguousarray(values, dtype=np.float64)
        c = np.ascontiguousarray(col_indices, dtype=np.uint64)
        r = np.ascontiguousarray(row_ptr, dtype=np.uint64)
        key = (id(values), id(col_indices), id(row_ptr), v.size, r.size, ncols)
        digest = hash((v.tobytes(), c.tobytes(), r.tobytes()))
        hit = self._resident.get(key)
        if hit is not None and hit[0] == digest:
            return hit[1]
        if hit is not None:
            hit[1].free()
        m = CsrMatrix.upload(v, c, r, ncols)
        self._resident[key] = (digest, m)
        return m

    def spmv(self, values, col_indices, row_ptr, x, y: Optional[list] = None) -> np.ndarray:
        self._require("spmv")
        x = np.ascontiguousarray(x, dtype=np.float64)
        nrows = max(len(row_ptr) - 1, 0)
        if len(values) == 0 or len(row_ptr) < 2:
            out = np.zeros(nrows)
        else:
            out = self._matrix(values, col_indices, row_ptr, x.size).spmv(x)
        if y is not None:
            y[:] = list(out)
        return out

    def spmm(self, values, col_indices, row_ptr, B, ncols_B: int, C_out: Optional[list] = None) -> np.ndarray:
        self._require("spmm")
        B = np.ascontiguousarray(B, dtype=np.float64).ravel()
        nrows = max(len(row_ptr) - 1, 0)
        if len(values) == 0 or len(row_ptr) < 2 or ncols_B == 0:
            out = np.zeros(nrows * ncols_B)
        else:
            ncols = B.size // ncols_B
            out = self._matrix(values, col_indices, row_ptr, ncols).spmm(B, ncols_B)
        if C_out is not None:
            C_out[:] = list(out)
        return out

    def dot(self, a, b) -> float:
        self._require("dot")
        a = np.ascontiguousarray(a, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        n = min(a.size, b.size)
        out = C.c_double()
        check(LIB.hn_dot(_dptr(a), _dptr(b), n, C.byref(out)))
        return out.value

    def axpy(self, alpha: float, x, y: np.ndarray) -> np.ndarray:
        self._require("axpy")
        x = np.ascontiguousarray(x, dtype=np.float64)
        if not (isinstance(y, np.ndarray) and y.dtype == np.float64 and y.flags.c_contiguous):
            raise TypeError("y must be a contiguous float64 numpy array (updated in place)")
        n = min(x.size, y.size)
        check(LIB.hn_axpy(float(alpha), _dptr(x), _dptr(y), n))
        return y

    def norm_inf(self, v) -> float:
        self._require("norm_inf")
        v = np.ascontiguousarray(v, dtype=np.float64)
        out = C.c_double()
        check(LIB.hn_norm_inf(_dptr(v), v.size, C.byref(out)))
        return out.value

    def device_memory(self) -> int:
        return mem_info()[1] if self._available else 0

    def free_memory(self) -> int:
        return mem_info()[0] if self._available else 0

    def was_gpu_kernel_executed(self) -> bool:
        return LIB.hn_kernel_executions() > 0

    def gpu_kernel_executions(self) -> int:
        return int(LIB.hn_kernel_executions())

    def reset_execution_stats(self) -> None:
        LIB.hn_reset_execution_stats()


def gpu_visible() -> bool:
    """True when a gfx950 device can be initialised."""
    info = DeviceInfo()
    return LIB.hn_init(-1, C.byref(info)) == HN_OK


# ---------------------------------------------------------------------------
# Interior-point solver (hn_ipm_solve)
# ---------------------------------------------------------------------------
HN_NE_AUTO, HN_NE_DENSE, HN_NE_PCG, HN_NE_SPARSE = 0, 1, 2, 3
HN_PREC_F64, HN_PREC_MIXED, HN_PREC_AUTO = 1, 2, 3
HN_PIVOT_CLAMP, HN_PIVOT_SKIP = 0, 1
HN_LOOP_GRAPH, HN_LOOP_DEVICE, HN_LOOP_HOST = 0, 1, 2
STATUS = ["UNKNOWN", "OPTIMAL", "INFEASIBLE", "UNBOUNDED", "SUBOPTIMAL", "TIME_LIMIT",
          "ITER_LIMIT", "NUMERICAL_ERROR", "INTERRUPTED"]

_i64p = C.POINTER(C.c_int64)
_i8p = C.POINTER(C.c_int8)
INTERRUPT_FN = C.CFUNCTYPE(C.c_int, C.c_void_p)


class _HnLp(C.Structure):
    _fields_ = [("m", C.c_int64), ("n", C.c_int64), ("row_ptr", _i64p), ("col", _i64p), ("val", _dp),
                ("sense", _i8p), ("rhs", _dp), ("lb", _dp), ("ub", _dp), ("obj", _dp), ("maximize", C.c_int),
                ("obj_offset", C.c_double), ("x0", _dp), ("y0", _dp), ("s0", _dp)]


class _HnOpts(C.Structure):
    _fields_ = [("max_iterations", C.c_int), ("time_limit_seconds", C.c_double),
                ("convergence_tol", C.c_double), ("complementarity_tol", C.c_double),
                ("regularization", C.c_double), ("interrupt_callback", INTERRUPT_FN),
                ("interrupt_user", C.c_void_p), ("solver", C.c_int), ("precision", C.c_int),
                ("mu_switch", C.c_double), ("dense_max_m", C.c_int64), ("pcg_max_iter", C.c_int),
                ("x0", _dp), ("y0", _dp), ("s0", _dp), ("verbose", C.c_int), ("pivot_rule", C.c_int),
                ("mu_target", C.c_double), ("gap_tol", C.c_double), ("endgame_mu", C.c_double),
                ("loop", C.c_int)]


class _HnResult(C.Structure):
    _fields_ = [("status", C.c_int), ("iterations", C.c_int64), ("objective", C.c_double),
                ("solve_ms", C.c_double), ("primal", _dp), ("x", _dp), ("y", _dp), ("s", _dp),
                ("cap_m", C.c_int64), ("cap_n", C.c_int64), ("sf_m", C.c_int64), ("sf_n", C.c_int64),
                ("rp_norm", C.c_double), ("rd_norm", C.c_double), ("mu", C.c_double),
                ("rel_primal", C.c_double), ("rel_dual", C.c_double), ("rel_gap", C.c_double),
                ("scaled_rel_primal", C.c_double), ("scaled_rel_dual", C.c_double),
                ("scaled_rel_gap", C.c_double), ("ne_solves", C.c_int64), ("refine_sweeps", C.c_int64), ("cg_iterations", C.c_int64),
                ("fp64_from_iter", C.c_int64), ("warm_started", C.c_int), ("projected", C.c_int),
                ("sparse_l_nnz", C.c_int64), ("device_steps", C.c_int64), ("host_redos", C.c_int64)]


LIB.hn_ipm_default_opts.argtypes = [C.POINTER(_HnOpts)]
LIB.hn_ipm_default_opts.restype = C.c_int
LIB.hn_ipm_solve.argtypes = [C.POINTER(_HnLp), C.POINTER(_HnOpts), C.POINTER(_HnResult)]
LIB.hn_ipm_solve.restype = C.c_int


class InteriorPointOptions:
    """Mirror of InteriorPointOptions (core/lp/interior_point.hpp:16-27) plus
    the GPU choices of hn_ipm_opts.  ``crossover`` must stay False: the
    simplex polish is the caller's (CPU) business."""

    def __init__(self, **kw):
        self.max_iterations = 100
        self.time_limit_seconds = 0.0
        self.convergence_tol = 1e-8
        self.complementarity_tol = 1e-8
        self.regularization = 1e-8
        self.crossover = False
        self.interrupt_callback = None
        self.solver = HN_NE_AUTO
        self.precision = HN_PREC_F64
        self.mu_switch = 1e-4
        self.dense_max_m = 40000
        self.pcg_max_iter = 2000
        self.warm_start = None          # (x0, y0, s0) of the scaled standard form
        self.pivot_rule = HN_PIVOT_CLAMP  # HN_PIVOT_SKIP: Wright's skipped pivots, r = regularization
        self.mu_target = 0.0            # > 0: return the first iterate with mu <= mu_target
        self.gap_tol = 0.0              # > 0: also stop only once the caller's LP has rel_gap <= gap_tol
        self.loop = HN_LOOP_GRAPH       # HN_LOOP_DEVICE / HN_LOOP_HOST: see hn_ipm_opts.loop
        self.verbose = False
        for k, v in kw.items():
            if not hasattr(self, k):
                raise TypeError(f"unknown option {k}")
            setattr(self, k, v)


class InteriorPointResult:
    def __init__(self, r: _HnResult, primal, x, y, s):
        self.status = STATUS[r.status]
        self.objective_value = r.objective
        self.iterations = int(r.iterations)
        self.solve_time_ms = r.solve_ms
        self.primal = primal
        self.x, self.y, self.s = x[:r.sf_n].copy(), y[:r.sf_m].copy(), s[:r.sf_n].copy()
        self.rp_norm, self.rd_norm, self.mu = r.rp_norm, r.rd_norm, r.mu
        # relative KKT measures on the caller's LP, and the same on the scaled standard form
        self.rel_primal, self.rel_dual, self.rel_gap = r.rel_primal, r.rel_dual, r.rel_gap
        self.scaled_rel_primal, self.scaled_rel_dual = r.scaled_rel_primal, r.scaled_rel_dual
        self.scaled_rel_gap = r.scaled_rel_gap
        self.ne_solves, self.refine_sweeps = int(r.ne_solves), int(r.refine_sweeps)
        self.cg_iterations, self.fp64_from_iter = int(r.cg_iterations), int(r.fp64_from_iter)
        self.warm_started = bool(r.warm_started)
        self.projected = bool(r.projected)
        self.sparse_l_nnz = int(r.sparse_l_nnz)
        self.device_steps, self.host_redos = int(r.device_steps), int(r.host_redos)

    def __repr__(self):
        return (f"InteriorPointResult({self.status}, obj={self.objective_value!r}, it={self.iterations}, "
                f"{self.solve_time_ms:.1f} ms)")


class InteriorPointSolver:
    """Mirror of hypernova::lp::InteriorPointSolver (interior_point.hpp:40-54)
    whose solve() runs hn_ipm_solve on the GPU."""

    def __init__(self, options: Optional[InteriorPointOptions] = None):
        self._opts = options or InteriorPointOptions()

    def options(self) -> InteriorPointOptions:
        return self._opts

    def set_options(self, o: InteriorPointOptions) -> None:
        self._opts = o

    def solve(self, lp) -> InteriorPointResult:
        o = self._opts
        if o.crossover:
            raise ValueError("crossover is not part of the GPU path; set crossover=False")
        keep = dict(rp=np.ascontiguousarray(lp.row_ptr, np.int64), ci=np.ascontiguousarray(lp.col, np.int64),
                    v=np.ascontiguousarray(lp.val, np.float64), sense=np.ascontiguousarray(lp.sense, np.int8),
                    rhs=np.ascontiguousarray(lp.rhs, np.float64), lb=np.ascontiguousarray(lp.lb, np.float64),
                    ub=np.ascontiguousarray(lp.ub, np.float64), obj=np.ascontiguousarray(lp.obj, np.float64))
        P = _HnLp(lp.m, lp.n, keep["rp"].ctypes.data_as(_i64p), keep["ci"].ctypes.data_as(_i64p),
                  _dptr(keep["v"]), keep["sense"].ctypes.data_as(_i8p), _dptr(keep["rhs"]), _dptr(keep["lb"]),
                  _dptr(keep["ub"]), _dptr(keep["obj"]), int(lp.maximize), float(getattr(lp, "obj_offset", 0.0)))
        opts = _HnOpts()
        check(LIB.hn_ipm_default_opts(C.byref(opts)))
        opts.max_iterations = o.max_iterations
        opts.time_limit_seconds = o.time_limit_seconds
        opts.convergence_tol = o.convergence_tol
        opts.complementarity_tol = o.complementarity_tol
        opts.regularization = o.regularization
        cb = None
        if o.interrupt_callback is not None:
            user_cb = o.interrupt_callback
            cb = INTERRUPT_FN(lambda _u: 1 if user_cb() else 0)
            opts.interrupt_callback = cb
        opts.solver = o.solver
        opts.precision = o.precision
        opts.mu_switch = o.mu_switch
        opts.dense_max_m = o.dense_max_m
        opts.pcg_max_iter = o.pcg_max_iter
        opts.verbose = int(o.verbose)
        opts.pivot_rule = o.pivot_rule
        opts.mu_target = o.mu_target
        opts.gap_tol = o.gap_tol
        opts.loop = o.loop
        if o.warm_start is not None:
            ws = [np.ascontiguousarray(a, np.float64) for a in o.warm_start]
            keep["ws"] = ws
            opts.x0, opts.y0, opts.s0 = (_dptr(a) for a in ws)
        cap_m = lp.m + lp.n + 8
        cap_n = 2 * lp.n + lp.m + 8
        primal = np.zeros(max(lp.n, 1))
        x, s, y = np.zeros(cap_n), np.zeros(cap_n), np.zeros(cap_m)
        r = _HnResult()
        r.primal, r.x, r.y, r.s = _dptr(primal), _dptr(x), _dptr(y), _dptr(s)
        r.cap_m, r.cap_n = cap_m, cap_n
        check(LIB.hn_ipm_solve(C.byref(P), C.byref(opts), C.byref(r)))
        return InteriorPointResult(r, primal[:lp.n].copy(), x, y, s)


def ipm_solve(lp, **kw) -> InteriorPointResult:
    return InteriorPointSolver(InteriorPointOptions(**kw)).solve(lp)


LIB.hn_ipm_solve_batch.argtypes = [C.POINTER(_HnLp), C.c_uint32, C.POINTER(_HnOpts), C.POINTER(_HnResult)]
LIB.hn_ipm_solve_batch.restype = C.c_int
LIB.hn_ipm_solve_batch_gnn.argtypes = [C.POINTER(_HnLp), C.c_uint32, C.POINTER(_HnOpts), C.c_void_p, C.c_void_p,
                                       C.POINTER(_HnResult)]
LIB.hn_ipm_solve_batch_gnn.restype = C.c_int


def ipm_solve_batch(lps, precision: int = HN_PREC_F64, max_iterations: int = 100, mu_switch: float = 1e-4,
                    want_points: bool = False, pivot_rule: int = HN_PIVOT_CLAMP, mu_target: float = 0.0,
                    warm=None, verbose: bool = False, gap_tol: float = 0.0, gnn=None, endgame_mu: float = 0.0):
    """Solves many LPs in one persistent launch (hn_ipm_solve_batch).
    warm: optional list (one entry per LP, or None) of (x0, y0, s0) warm
    starts of the scaled standard forms.  verbose: the library prints the
    time of each phase of the call on stderr; verbose=2 also times the
    phases inside the kernel (start, residuals, projection, right-hand
    side, assembly, diagonal blocks, panels, trailing updates, triangular
    solves, refinement, step) and prints each one's share of the
    workgroups' time.  gap_tol > 0: an instance
    stops only once its relative gap |c'x - b'y| / (1 + |c'x|) is also at
    most gap_tol (besides the reference's absolute test).  gnn: optional
    (network handle, graph handle) - LPs without a warm start whose standard
    form has the graph's pattern are warm-started by the network on the
    device inside the call (hn_ipm_solve_batch_gnn).  Returns one
    InteriorPointResult per LP (solve_time_ms = whole call)."""
    count = len(lps)
    arr = (_HnLp * count)()
    keep = []
    for k, lp in enumerate(lps):
        d = dict(rp=np.ascontiguousarray(lp.row_ptr, np.int64), ci=np.ascontiguousarray(lp.col, np.int64),
                 v=np.ascontiguousarray(lp.val, np.float64), sense=np.ascontiguousarray(lp.sense, np.int8),
                 rhs=np.ascontiguousarray(lp.rhs, np.float64), lb=np.ascontiguousarray(lp.lb, np.float64),
                 ub=np.ascontiguousarray(lp.ub, np.float64), obj=np.ascontiguousarray(lp.obj, np.float64))
        keep.append(d)
        arr[k] = _HnLp(lp.m, lp.n, d["rp"].ctypes.data_as(_i64p), d["ci"].ctypes.data_as(_i64p), _dptr(d["v"]),
                       d["sense"].ctypes.data_as(_i8p), _dptr(d["rhs"]), _dptr(d["lb"]), _dptr(d["ub"]),
                       _dptr(d["obj"]), int(lp.maximize), float(getattr(lp, "obj_offset", 0.0)))
        if warm is not None and warm[k] is not None:
            d["warm"] = [np.ascontiguousarray(a, np.float64) for a in warm[k]]
            arr[k].x0, arr[k].y0, arr[k].s0 = (_dptr(a) for a in d["warm"])
    opts = _HnOpts()
    check(LIB.hn_ipm_default_opts(C.byref(opts)))
    opts.precision = precision
    opts.max_iterations = max_iterations
    opts.mu_switch = mu_switch
    opts.pivot_rule = pivot_rule
    opts.mu_target = mu_target
    opts.gap_tol = gap_tol
    opts.endgame_mu = endgame_mu
    opts.verbose = int(verbose)
    res = (_HnResult * count)()
    bufs = []
    for k, lp in enumerate(lps):
        primal = np.zeros(max(lp.n, 1))
        cap_m, cap_n = (lp.m + lp.n + 8, 2 * lp.n + lp.m + 8) if want_points else (0, 0)
        x, y, s = np.zeros(max(cap_n, 1)), np.zeros(max(cap_m, 1)), np.zeros(max(cap_n, 1))
        bufs.append((primal, x, y, s))
        res[k].primal = _dptr(primal)
        if want_points:
            res[k].x, res[k].y, res[k].s = _dptr(x), _dptr(y), _dptr(s)
            res[k].cap_m, res[k].cap_n = cap_m, cap_n
    if gnn is not None:
        check(LIB.hn_ipm_solve_batch_gnn(arr, count, C.byref(opts), gnn[0], gnn[1], res))
    else:
        check(LIB.hn_ipm_solve_batch(arr, count, C.byref(opts), res))
    return [InteriorPointResult(res[k], bufs[k][0][:lps[k].n].copy(), bufs[k][1], bufs[k][2], bufs[k][3])
            for k in range(count)]


HN_TUNE_BUCKETS = 6


class _HnTuning(C.Structure):
    _fields_ = [("m_hi", C.c_int * HN_TUNE_BUCKETS), ("lps", C.c_int * HN_TUNE_BUCKETS),
                ("precision", C.c_int * HN_TUNE_BUCKETS), ("ms_f64", C.c_double * HN_TUNE_BUCKETS),
                ("ms_mixed", C.c_double * HN_TUNE_BUCKETS)]


LIB.hn_batch_autotune.argtypes = [C.POINTER(_HnLp), C.c_uint32, C.POINTER(_HnOpts), C.c_uint32, C.POINTER(_HnTuning)]
LIB.hn_batch_autotune.restype = C.c_int


def batch_autotune(lps, per_bucket: int = 128, pivot_rule: int = HN_PIVOT_CLAMP, gap_tol: float = 0.0,
                   endgame_mu: float = 0.0, max_iterations: int = 100, mu_switch: float = 1e-4) -> list:
    """Autotunes the batched solver on a sample (hn_batch_autotune): per m
    bucket of the standard form, the batch of at most per_bucket sample LPs
    is solved in FP64 and with the FP32 factor, and the faster precision is
    stored for the device (ipm_solve_batch(precision=HN_PREC_AUTO) uses it).
    Returns one dict per bucket: m_hi, lps, precision, ms_f64, ms_mixed."""
    count = len(lps)
    arr = (_HnLp * count)()
    keep = []
    for k, lp in enumerate(lps):
        d = dict(rp=np.ascontiguousarray(lp.row_ptr, np.int64), ci=np.ascontiguousarray(lp.col, np.int64),
                 v=np.ascontiguousarray(lp.val, np.float64), sense=np.ascontiguousarray(lp.sense, np.int8),
                 rhs=np.ascontiguousarray(lp.rhs, np.float64), lb=np.ascontiguousarray(lp.lb, np.float64),
                 ub=np.ascontiguousarray(lp.ub, np.float64), obj=np.ascontiguousarray(lp.obj, np.float64))
        keep.append(d)
        arr[k] = _HnLp(lp.m, lp.n, d["rp"].ctypes.data_as(_i64p), d["ci"].ctypes.data_as(_i64p), _dptr(d["v"]),
                       d["sense"].ctypes.data_as(_i8p), _dptr(d["rhs"]), _dptr(d["lb"]), _dptr(d["ub"]),
                       _dptr(d["obj"]), int(lp.maximize), float(getattr(lp, "obj_offset", 0.0)))
    opts = _HnOpts()
    check(LIB.hn_ipm_default_opts(C.byref(opts)))
    opts.pivot_rule = pivot_rule
    opts.gap_tol = gap_tol
    opts.endgame_mu = endgame_mu
    opts.max_iterations = max_iterations
    opts.mu_switch = mu_switch
    t = _HnTuning()
    check(LIB.hn_batch_autotune(arr, count, C.byref(opts), per_bucket, C.byref(t)))
    return [{"m_hi": t.m_hi[b], "lps": t.lps[b],
             "precision": {HN_PREC_F64: "f64", HN_PREC_MIXED: "mixed"}[t.precision[b]],
             "ms_f64": round(t.ms_f64[b], 2), "ms_mixed": round(t.ms_mixed[b], 2)} for b in range(HN_TUNE_BUCKETS)]


LIB.hn_lp_standard_form.argtypes = [C.POINTER(_HnLp), C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                    C.POINTER(C.c_int64), _i64p, _i64p, _dp, _dp, _dp]
LIB.hn_lp_standard_form.restype = C.c_int


def lp_standard_form(lp) -> dict:
    """The scaled standard form (M as CSR, rhs, c) the IPM iterates on
    (hn_lp_standard_form); host-side, no device needed."""
    d = dict(rp=np.ascontiguousarray(lp.row_ptr, np.int64), ci=np.ascontiguousarray(lp.col, np.int64),
             v=np.ascontiguousarray(lp.val, np.float64), sense=np.ascontiguousarray(lp.sense, np.int8),
             rhs=np.ascontiguousarray(lp.rhs, np.float64), lb=np.ascontiguousarray(lp.lb, np.float64),
             ub=np.ascontiguousarray(lp.ub, np.float64), obj=np.ascontiguousarray(lp.obj, np.float64))
    P = _HnLp(lp.m, lp.n, d["rp"].ctypes.data_as(_i64p), d["ci"].ctypes.data_as(_i64p), _dptr(d["v"]),
              d["sense"].ctypes.data_as(_i8p), _dptr(d["rhs"]), _dptr(d["lb"]), _dptr(d["ub"]), _dptr(d["obj"]),
              int(lp.maximize), float(getattr(lp, "obj_offset", 0.0)))
    # one call into buffers of guaranteed capacity: every kept row gives one
    # row and every boxed column one more; a column gives at most two slots
    # plus a slack of its bound row, a row at most one slack
    m0, n0, z0 = int(lp.m), int(lp.n), int(d["rp"][-1]) if lp.m else 0
    cap_m, cap_n, cap_z = m0 + n0, 3 * n0 + m0, 2 * z0 + m0 + 2 * n0
    rp = np.empty(cap_m + 1, np.int64)
    ci = np.empty(max(cap_z, 1), np.int64)
    v = np.empty(max(cap_z, 1))
    rhs = np.empty(max(cap_m, 1))
    c = np.empty(max(cap_n, 1))
    sm, sn, nz = C.c_int64(), C.c_int64(), C.c_int64()
    check(LIB.hn_lp_standard_form(C.byref(P), C.byref(sm), C.byref(sn), C.byref(nz), rp.ctypes.data_as(_i64p),
                                  ci.ctypes.data_as(_i64p), _dptr(v), _dptr(rhs), _dptr(c)))
    m, n, z = sm.value, sn.value, nz.value
    return dict(m=m, n=n, rp=rp[:m + 1].copy(), ci=ci[:z].copy(), v=v[:z].copy(), rhs=rhs[:m].copy(),
                c=c[:n].copy())


# ---- post-solve verification (hn_verify) ----------------------------------------
class _HnTol(C.Structure):
    _fields_ = [("feasibility_tol", C.c_double), ("optimality_tol", C.c_double),
                ("complementarity_tol", C.c_double), ("integrality_tol", C.c_double)]


class _HnSolution(C.Structure):
    _fields_ = [("primal", _dp), ("n_primal", C.c_uint64), ("dual", _dp), ("n_dual", C.c_uint64),
                ("reduced_costs", _dp), ("n_reduced_costs", C.c_uint64), ("objective_value", C.c_double)]


class _HnVerifyResult(C.Structure):
    _fields_ = [("feasible", C.c_int), ("optimal", C.c_int), ("primal_infeasibility", C.c_double),
                ("dual_infeasibility", C.c_double), ("complementarity", C.c_double),
                ("integrality_violation", C.c_double), ("objective_discrepancy", C.c_double),
                ("n_violated", C.c_uint64), ("violated", _u64p), ("violated_capacity", C.c_uint64),
                ("message", C.c_char * 160), ("device_ms", C.c_double)]


LIB.hn_default_tolerances.argtypes = [C.POINTER(_HnTol)]
LIB.hn_default_tolerances.restype = C.c_int
LIB.hn_verify.argtypes = [C.POINTER(_HnLp), _i8p, C.POINTER(_HnSolution), C.POINTER(_HnTol),
                          C.POINTER(_HnVerifyResult)]
LIB.hn_verify.restype = C.c_int

HN_VAR_CONTINUOUS, HN_VAR_INTEGER, HN_VAR_BINARY = 0, 1, 2


@dataclass
class ToleranceConfig:
    """The fields of numerical::ToleranceConfig the verifier reads, at their
    industrial_defaults (core/numerical/tolerance.hpp:60-75)."""
    feasibility_tol: float = 1e-9
    optimality_tol: float = 1e-9
    complementarity_tol: float = 1e-6
    integrality_tol: float = 1e-6


@dataclass
class Solution:
    """model::Solution as the verifier reads it (core/model/problem.hpp:172-190);
    empty dual / reduced_costs are absent, as in the reference."""
    primal: np.ndarray
    dual: Optional[np.ndarray] = None
    reduced_costs: Optional[np.ndarray] = None
    objective_value: float = 0.0


@dataclass
class VerificationResult:
    """validation::VerificationResult (core/validation/solution_verifier.hpp:12-21)."""
    feasible: bool
    optimal: bool
    primal_infeasibility: float
    dual_infeasibility: float
    complementarity: float
    integrality_violation: float
    objective_discrepancy: float
    violated_constraints: list
    n_violated: int
    message: str
    device_ms: float = 0.0


class SolutionVerifier:
    """SolutionVerifier (core/validation/solution_verifier.hpp:23-41) on the
    device through hn_verify.  var_type (model::VarType codes, one per
    column) marks integer columns; None means all continuous."""

    def __init__(self, tol: Optional[ToleranceConfig] = None, max_listed: int = 1 << 16):
        self.tol = tol or ToleranceConfig()
        self.max_listed = max_listed

    def verify(self, lp, solution: Solution, var_type=None) -> bool:
        r = self.verify_detailed(lp, solution, var_type)
        return r.feasible and r.optimal

    def verify_feasible(self, lp, solution: Solution, var_type=None) -> bool:
        return self.verify_detailed(lp, solution, var_type).feasible

    def verify_detailed(self, lp, solution: Solution, var_type=None) -> VerificationResult:
        keep = dict(rp=np.ascontiguousarray(lp.row_ptr, np.int64), ci=np.ascontiguousarray(lp.col, np.int64),
                    v=np.ascontiguousarray(lp.val, np.float64), sense=np.ascontiguousarray(lp.sense, np.int8),
                    rhs=np.ascontiguousarray(lp.rhs, np.float64), lb=np.ascontiguousarray(lp.lb, np.float64),
                    ub=np.ascontiguousarray(lp.ub, np.float64), obj=np.ascontiguousarray(lp.obj, np.float64))
        P = _HnLp(lp.m, lp.n, keep["rp"].ctypes.data_as(_i64p), keep["ci"].ctypes.data_as(_i64p),
                  _dptr(keep["v"]), keep["sense"].ctypes.data_as(_i8p), _dptr(keep["rhs"]), _dptr(keep["lb"]),
                  _dptr(keep["ub"]), _dptr(keep["obj"]), int(lp.maximize), float(getattr(lp, "obj_offset", 0.0)))
        vec = {}
        for name in ("primal", "dual", "reduced_costs"):
            a = getattr(solution, name)
            vec[name] = None if a is None else np.ascontiguousarray(a, np.float64)
        S = _HnSolution()
        for name in ("primal", "dual", "reduced_costs"):
            a = vec[name]
            setattr(S, name, _dptr(a) if a is not None and a.size else None)
        S.n_primal = 0 if vec["primal"] is None else vec["primal"].size
        S.n_dual = 0 if vec["dual"] is None else vec["dual"].size
        S.n_reduced_costs = 0 if vec["reduced_costs"] is None else vec["reduced_costs"].size
        S.objective_value = float(solution.objective_value)
        t = _HnTol(self.tol.feasibility_tol, self.tol.optimality_tol, self.tol.complementarity_tol,
                   self.tol.integrality_tol)
        vt = None if var_type is None else np.ascontiguousarray(var_type, np.int8)
        listed = np.zeros(max(self.max_listed, 1), np.uint64)
        out = _HnVerifyResult()
        out.violated = listed.ctypes.data_as(_u64p)
        out.violated_capacity = self.max_listed
        check(LIB.hn_verify(C.byref(P), None if vt is None else vt.ctypes.data_as(_i8p), C.byref(S), C.byref(t),
                            C.byref(out)))
        k = min(int(out.n_violated), self.max_listed)
        return VerificationResult(bool(out.feasible), bool(out.optimal), out.primal_infeasibility,
                                  out.dual_infeasibility, out.complementarity, out.integrality_violation,
                                  out.objective_discrepancy, [int(q) for q in listed[:k]], int(out.n_violated),
                                  out.message.decode(), out.device_ms)


def verify(lp, primal, dual=None, reduced_costs=None, objective_value: float = 0.0, var_type=None,
           tol: Optional[ToleranceConfig] = None) -> VerificationResult:
    return SolutionVerifier(tol).verify_detailed(lp, Solution(primal, dual, reduced_costs, objective_value),
                                                 var_type)
