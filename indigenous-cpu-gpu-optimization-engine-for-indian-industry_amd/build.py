"""Builds the native libraries of the package in-tree.

* ``lib/libhn_kernels.so`` - the HIP kernel library for gfx950 (hipcc), the
  C-ABI declared in ``include/hn_kernels.h``.
* ``lib/libhn_gen.so`` - the seeded synthetic-input generators (host C++).
* ``lib/libhn_backend.so`` - the C++ ``HipBackend`` plugin that mirrors the
  reference's ``IComputeBackend`` on top of ``libhn_kernels.so``.

Run ``python build.py`` or call :func:`build_all`.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
OBJ = os.path.join(HERE, "_build")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

# Packed FP32 instructions are disabled for all device code.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + INCLUDE, "-I" + CSRC,
    "-Wno-unused-result",
]

KERNEL_SOURCES = [
    "hn_runtime.hip",
    "hn_spmv.hip",
    "hn_spmv_slab.hip",
    "hn_transpose.hip",
    "hn_vec.hip",
    "hn_stdform.cpp",
    "hn_normal.hip",
    "hn_ipm.hip",
    "hn_batch.hip",
    "hn_gnn.hip",
    "hn_verify.hip",
    "hn_order.cpp",
    "hn_sparse_ldl.hip",
]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{out}")


def _stale(target: str, deps: list[str]) -> bool:
    if not os.path.exists(target):
        return True
    t = os.path.getmtime(target)
    return any(os.path.getmtime(d) > t for d in deps)


def _headers() -> list[str]:
    hs = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hpp", ".h"))]
    hs += [os.path.join(INCLUDE, f) for f in os.listdir(INCLUDE) if f.endswith(".h")]
    return hs


def build_kernels(force: bool = True, jobs: int = 8) -> str:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)
    headers = _headers()
    objs = []
    cmds = []
    for src in KERNEL_SOURCES:
        s = os.path.join(CSRC, src)
        o = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
        objs.append(o)
        if force or _stale(o, [s] + headers):
            cmds.append([HIPCC, *HIP_FLAGS, "-c", s, "-o", o])
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    so = os.path.join(LIB, "libhn_kernels.so")
    tmp = so + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, so)
    return so


def build_gen() -> str:
    os.makedirs(LIB, exist_ok=True)
    so = os.path.join(LIB, "libhn_gen.so")
    tmp = so + ".tmp"
    _run([CXX, "-O2", "-std=c++17", "-fPIC", "-shared", "-I" + INCLUDE,
          os.path.join(CSRC, "hn_gen.cpp"), "-o", tmp])
    os.replace(tmp, so)
    return so


def build_backend() -> list[str]:
    """libhn_backend.so (the C++ HipBackend plugin) and its test program."""
    os.makedirs(LIB, exist_ok=True)
    so = os.path.join(LIB, "libhn_backend.so")
    _run([CXX, "-O2", "-std=c++17", "-fPIC", "-shared", "-I" + CSRC, os.path.join(CSRC, "hip_backend.cpp"),
          "-o", so + ".tmp", "-ldl"])
    os.replace(so + ".tmp", so)
    exe = os.path.join(LIB, "test_hip_backend")
    _run([CXX, "-O2", "-std=c++17", "-I" + CSRC, os.path.join(ROOT, "tests", "cpp", "test_hip_backend.cpp"),
          os.path.join(CSRC, "hip_backend.cpp"), "-o", exe + ".tmp", "-ldl"])
    os.replace(exe + ".tmp", exe)
    return [so, exe]


def build_all(force: bool = True) -> list[str]:
    return [build_kernels(force=force), build_gen(), *build_backend()]


if __name__ == "__main__":
    for p in build_all(force="--incremental" not in sys.argv):
        print(p)
