"""Writes tests/golden/backend_cases.npz: the inputs of the reference's backend
tests (regenerated by mt_cases.cpp with their std::mt19937 recipe) and the
outputs of the reference's own CPU backend on them (oracle/_ref).

Run from the repository root in a container that has the reference build:
    python tests/golden/make_backend_fixtures.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402

# name: (nrows, ncols, density, seed, [k, Bseed, Blo])  -- test_gpu_backend.cpp line
CASES = {
    "cpu_spmv": (200, 150, 0.10, 42),            # :90-98
    "cpu_spmm": (80, 60, 0.2, 7, 4, 99, -2.0),   # :100-124
    "auto_spmv": (64, 48, 0.3, 5),               # :163-175
    "cuda_scalar": (1500, 1200, 0.003, 77),      # :211-222
    "cuda_warp": (300, 800, 0.08, 4242),         # :224-234
    "cuda_spmm": (400, 300, 0.05, 2024, 5, 8, -1.0),  # :260-276
    "residency": (2000, 1500, 0.03, 555),        # :278-292
}


def main():
    exe = os.path.join(tempfile.mkdtemp(), "mt_cases")
    subprocess.run(["g++", "-O2", "-std=c++17", os.path.join(HERE, "mt_cases.cpp"), "-o", exe], check=True)
    out = {}
    for name, spec in CASES.items():
        nrows, ncols, dens, seed = spec[:4]
        args = [exe, str(nrows), str(ncols), repr(dens), str(seed)] + [repr(a) if isinstance(a, float) else str(a)
                                                                      for a in spec[4:]]
        raw = subprocess.run(args, check=True, stdout=subprocess.PIPE).stdout
        nnz = int(np.frombuffer(raw[:8], np.int64)[0])
        off = 8
        rp = np.frombuffer(raw[off:off + 8 * (nrows + 1)], np.int64); off += 8 * (nrows + 1)
        ci = np.frombuffer(raw[off:off + 8 * nnz], np.int64); off += 8 * nnz
        v = np.frombuffer(raw[off:off + 8 * nnz], np.float64); off += 8 * nnz
        x = np.frombuffer(raw[off:off + 8 * ncols], np.float64); off += 8 * ncols
        out[f"{name}_shape"] = np.array([nrows, ncols], np.int64)
        out[f"{name}_rp"] = rp.astype(np.int32)
        out[f"{name}_ci"] = ci.astype(np.int32)
        out[f"{name}_v"] = v
        out[f"{name}_x"] = x
        out[f"{name}_y"] = O.ref_cpu_spmv(rp, ci, v, x)
        if len(spec) > 4:
            k = spec[4]
            B = np.frombuffer(raw[off:off + 8 * ncols * k], np.float64)
            out[f"{name}_B"] = B
            out[f"{name}_k"] = np.array([k], np.int64)
    np.savez_compressed(os.path.join(HERE, "backend_cases.npz"), **out)
    print({k: v.shape for k, v in out.items() if k.endswith("_v")})


if __name__ == "__main__":
    main()
