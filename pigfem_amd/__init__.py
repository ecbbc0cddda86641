"""pigfem_amd -- MI355X-native K/R element assembly for PIGFEM-style IGFEM RVEs.

The package holds only the assembly hot path: HIP kernels and their C ABI
(csrc/, include/), the host set-up harness that produces the problem
description, and thin Python mirrors of the reference's Assembler API.
"""
from ._ffi import PG_FUNC, PG_JAC, PG_FP32, PG_BATCHED, PgError, lib  # noqa: F401
from .problem import RVEConfig, RVEProblem, batch_config, config, small_config  # noqa: F401
from .assembler import AssemblerGPU, AssemblerGPUThermal, AssemblerGPUThermalPreLoad, PMatrix, PVector, cut_inclusions, write_cohesive_openings, \
    write_cohesive_failure_lengths  # noqa: F401
from .device import DeviceBuffer, synchronize  # noqa: F401
from .sensitivity import SensitivityMaterialParameterGPU, lei_dmat_sensitivity  # noqa: F401
from .vtk import write_vtk  # noqa: F401

__all__ = ["RVEConfig", "RVEProblem", "batch_config", "config", "small_config", "AssemblerGPU",
           "AssemblerGPUThermal", "AssemblerGPUThermalPreLoad", "PMatrix", "PVector",
           "cut_inclusions", "write_cohesive_openings", "write_cohesive_failure_lengths",
           "DeviceBuffer", "synchronize", "SensitivityMaterialParameterGPU", "lei_dmat_sensitivity", "write_vtk", "PG_FUNC", "PG_JAC", "PG_FP32", "PG_BATCHED", "PgError", "lib"]
