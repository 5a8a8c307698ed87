"""Solver-level choices around the GPU path, as hypernova::Solver::solve makes
them (api/cpp/api.cpp), with the HIP backend in place of the CUDA one:

* which compute backend a solve gets from its options (api.cpp:110-117),
* which LP engine the AUTO dispatch picks (determine_lp_engine,
  api.cpp:143-168): the GPU interior-point path is reached only for large
  sparse LPs or when the caller asks for the interior-point engine,
* the backend_used string of the result (api.cpp:529-547).

The rules and thresholds are the reference's; only "cuda" becomes "hip".
"""
from __future__ import annotations

import enum
from dataclasses import dataclass


class ComputeTarget(enum.IntEnum):
    """api.hpp ComputeTarget."""
    CPU_ONLY = 0
    CPU_GPU_AUTO = 1
    CPU_GPU_FORCE = 2


class EngineType(enum.IntEnum):
    """api.hpp EngineType."""
    AUTO = 0
    PRIMAL_SIMPLEX = 1
    DUAL_SIMPLEX = 2
    INTERIOR_POINT = 3
    BRANCH_AND_BOUND = 4
    BRANCH_AND_CUT = 5
    QP_ACTIVE_SET = 6
    QP_INTERIOR_POINT = 7


class BackendType(enum.IntEnum):
    """The factory's backend kinds (gpu_backend.hpp ComputeBackendType), with
    HIP for the reference's CUDA entry."""
    CPU = 0
    HIP = 1
    AUTO = 2


@dataclass
class SolverOptions:
    """The fields of api.hpp SolverOptions these choices read."""
    engine: EngineType = EngineType.AUTO
    compute_target: ComputeTarget = ComputeTarget.CPU_GPU_AUTO
    use_gpu: bool = True


def backend_type(options: SolverOptions) -> BackendType:
    """The backend a solve is given (api.cpp:110-117)."""
    if not options.use_gpu or options.compute_target == ComputeTarget.CPU_ONLY:
        return BackendType.CPU
    if options.compute_target == ComputeTarget.CPU_GPU_FORCE:
        return BackendType.HIP
    return BackendType.AUTO


def determine_lp_engine(n_vars: int, n_cons: int, nnz: int,
                        engine: EngineType = EngineType.AUTO) -> EngineType:
    """The AUTO dispatch for an LP (api.cpp:143-168): density = nnz / (n m);
    fewer than 500 variables -> primal simplex; density above 0.20 -> primal
    simplex; at least 10,000 variables and density below 0.05 -> interior
    point; otherwise primal simplex."""
    if engine != EngineType.AUTO:
        return engine
    density = 0.0
    if n_vars > 0 and n_cons > 0:
        density = float(nnz) / (float(n_vars) * float(n_cons))
    if n_vars < 500:
        return EngineType.PRIMAL_SIMPLEX
    if density > 0.20:
        return EngineType.PRIMAL_SIMPLEX
    if n_vars >= 10000 and density < 0.05:
        return EngineType.INTERIOR_POINT
    return EngineType.PRIMAL_SIMPLEX


def backend_used(options: SolverOptions, gpu_kernel_executed: bool, hip_available: bool) -> str:
    """The result's backend_used (api.cpp:529-547)."""
    if not options.use_gpu or options.compute_target == ComputeTarget.CPU_ONLY:
        return "cpu (GPU disabled)"
    if gpu_kernel_executed:
        return "hip (gfx950 kernels executed)"
    if options.compute_target == ComputeTarget.CPU_GPU_FORCE:
        return "hip (forced)" if hip_available else "cpu (hip forced but unavailable - fallback)"
    return "cpu (cost model selected CPU for matrix size)" if hip_available else "cpu (hip unavailable)"
