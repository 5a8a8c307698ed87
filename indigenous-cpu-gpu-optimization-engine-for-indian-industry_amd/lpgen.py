"""Seeded synthetic inputs (ctypes over ``libhn_gen.so``; see ``csrc/hn_gen.h``).

Workloads (sizes as used by the benchmarks and tests):

* ``random_lp(seed, m, n, k=5)`` - standard-form LP  min c'x, Ax=b, x>=0,
  k distinct rows per column, N(0,1) values, feasible and bounded by
  construction (x0, y0, s0 returned as a certificate).  2,000 x 5,000 is the
  single-LP workload.
* ``powerlaw_lp(seed, m, n, alpha, cmin, cmax)`` - 200,000 x 500,000 with
  column counts on [2, 200] is the large mixed-precision LP.
* ``batch_lp(i)`` - instance i of the 4,096-LP batch (seed 1000+i, m
  log-uniform in [50, 2000], n = 2.5 m).
* ``transport_lp(seed, S=200, T=300)`` - the transportation family.
* ``powerlaw_spmv(seed, nrows, ncols, nnz, s, cap)`` - the SpMV matrix
  (2,000,000 x 5,000,000, 40M nonzeros, Zipf 1.2 rows capped at 10,000).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GEN_LIB = os.path.join(HERE, "lib", "libhn_gen.so")

LE, GE, EQ = 0, 1, 2


class _GenLP(C.Structure):
    _fields_ = [("m", C.c_int64), ("n", C.c_int64), ("nnz", C.c_int64),
                ("row_ptr", C.POINTER(C.c_int64)), ("col", C.POINTER(C.c_int64)),
                ("val", C.POINTER(C.c_double)), ("rhs", C.POINTER(C.c_double)),
                ("sense", C.POINTER(C.c_int8)), ("lb", C.POINTER(C.c_double)),
                ("ub", C.POINTER(C.c_double)), ("obj", C.POINTER(C.c_double)),
                ("maximize", C.c_int), ("x0", C.POINTER(C.c_double)),
                ("y0", C.POINTER(C.c_double)), ("s0", C.POINTER(C.c_double))]


class _GenSpmv(C.Structure):
    _fields_ = [("nrows", C.c_int64), ("ncols", C.c_int64), ("nnz", C.c_int64),
                ("row_ptr", C.POINTER(C.c_int64)), ("col", C.POINTER(C.c_int32)),
                ("val", C.POINTER(C.c_double)), ("x", C.POINTER(C.c_double)),
                ("w", C.POINTER(C.c_double))]


def _load():
    if not os.path.exists(GEN_LIB):
        raise ImportError(f"{GEN_LIB} is not built; run build.py")
    lib = C.CDLL(GEN_LIB)
    lib.hn_gen_random_lp.restype = C.POINTER(_GenLP)
    lib.hn_gen_random_lp.argtypes = [C.c_uint64, C.c_int64, C.c_int64, C.c_int]
    lib.hn_gen_powerlaw_lp.restype = C.POINTER(_GenLP)
    lib.hn_gen_powerlaw_lp.argtypes = [C.c_uint64, C.c_int64, C.c_int64, C.c_double, C.c_int, C.c_int]
    lib.hn_gen_batch_lp.restype = C.POINTER(_GenLP)
    lib.hn_gen_batch_lp.argtypes = [C.c_uint64, C.c_int64, C.c_int64, C.c_int]
    lib.hn_gen_transport_lp.restype = C.POINTER(_GenLP)
    lib.hn_gen_transport_lp.argtypes = [C.c_uint64, C.c_int, C.c_int]
    lib.hn_gen_lp_free.argtypes = [C.POINTER(_GenLP)]
    lib.hn_gen_powerlaw_spmv.restype = C.POINTER(_GenSpmv)
    lib.hn_gen_powerlaw_spmv.argtypes = [C.c_uint64, C.c_int64, C.c_int64, C.c_int64, C.c_double, C.c_int64]
    lib.hn_gen_spmv_free.argtypes = [C.POINTER(_GenSpmv)]
    lib.hn_gen_uniform.argtypes = [C.c_uint64, C.c_double, C.c_double, C.c_int64, C.POINTER(C.c_double)]
    return lib


LIB = _load()


@dataclass
class LP:
    """A general LP in the reference's Problem shape (core/model/problem.hpp:79-170):
    CSR constraint matrix, one sense and rhs per row, bounds and objective per column."""
    m: int
    n: int
    row_ptr: np.ndarray
    col: np.ndarray
    val: np.ndarray
    rhs: np.ndarray
    sense: np.ndarray
    lb: np.ndarray
    ub: np.ndarray
    obj: np.ndarray
    maximize: bool = False
    obj_offset: float = 0.0
    x0: Optional[np.ndarray] = None
    y0: Optional[np.ndarray] = None
    s0: Optional[np.ndarray] = None

    @property
    def nnz(self) -> int:
        return int(self.val.size)


def _arr(ptr, n, dtype):
    if not ptr or n == 0:
        return np.zeros(0, dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


def _take_lp(p) -> LP:
    if not p:
        raise ValueError("generator rejected its arguments")
    g = p.contents
    m, n, z = g.m, g.n, g.nnz
    lp = LP(m=m, n=n, row_ptr=_arr(g.row_ptr, m + 1, np.int64), col=_arr(g.col, z, np.int64),
            val=_arr(g.val, z, np.float64), rhs=_arr(g.rhs, m, np.float64),
            sense=_arr(g.sense, m, np.int8), lb=_arr(g.lb, n, np.float64), ub=_arr(g.ub, n, np.float64),
            obj=_arr(g.obj, n, np.float64), maximize=bool(g.maximize),
            x0=_arr(g.x0, n, np.float64) if g.x0 else None,
            y0=_arr(g.y0, m, np.float64) if g.y0 else None,
            s0=_arr(g.s0, n, np.float64) if g.s0 else None)
    LIB.hn_gen_lp_free(p)
    return lp


def random_lp(seed: int, m: int, n: int, k: int = 5) -> LP:
    return _take_lp(LIB.hn_gen_random_lp(seed, m, n, k))


def powerlaw_lp(seed: int, m: int, n: int, alpha: float = 2.179, cmin: int = 2, cmax: int = 200) -> LP:
    return _take_lp(LIB.hn_gen_powerlaw_lp(seed, m, n, alpha, cmin, cmax))


def batch_lp(i: int, mlo: int = 50, mhi: int = 2000, k: int = 5) -> LP:
    return _take_lp(LIB.hn_gen_batch_lp(1000 + i, mlo, mhi, k))


def transport_lp(seed: int, S: int = 200, T: int = 300) -> LP:
    return _take_lp(LIB.hn_gen_transport_lp(seed, S, T))


@dataclass
class SpmvCase:
    nrows: int
    ncols: int
    row_ptr: np.ndarray
    col: np.ndarray
    val: np.ndarray
    x: np.ndarray
    w: np.ndarray

    @property
    def nnz(self) -> int:
        return int(self.val.size)


def powerlaw_spmv(seed: int = 2, nrows: int = 2_000_000, ncols: int = 5_000_000, nnz: int = 40_000_000,
                  s: float = 1.2, cap: int = 10_000) -> SpmvCase:
    p = LIB.hn_gen_powerlaw_spmv(seed, nrows, ncols, nnz, s, cap)
    if not p:
        raise ValueError("generator rejected its arguments")
    g = p.contents
    case = SpmvCase(nrows=g.nrows, ncols=g.ncols, row_ptr=_arr(g.row_ptr, g.nrows + 1, np.int64),
                    col=_arr(g.col, g.nnz, np.int32), val=_arr(g.val, g.nnz, np.float64),
                    x=_arr(g.x, g.ncols, np.float64), w=_arr(g.w, g.nrows, np.float64))
    LIB.hn_gen_spmv_free(p)
    return case


def uniform(seed: int, lo: float, hi: float, n: int) -> np.ndarray:
    out = np.empty(n, np.float64)
    LIB.hn_gen_uniform(seed, lo, hi, n, out.ctypes.data_as(C.POINTER(C.c_double)))
    return out
