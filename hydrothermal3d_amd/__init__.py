"""hydrothermal3d_amd -- MI355X (gfx950) HIP implementation of the
HydroThermal3D solver hot path (DYREL pseudo-transient Darcy/heat solves,
Gershgorin pseudo-time steps, residual norms and WENO5 advection).

Layers:
  kernels   reference-named per-kernel entry points (drop-in for @parallel launches)
  solver    fused DYREL solve and the closed-box physical step (HydroThermal3D)
  inputs    workloads and seeded synthetic inputs
  device    DeviceArray and the host reductions
The HIP library is hydrothermal3d_amd/libht3d.so (C ABI: include/ht3d.h).
"""
from . import _lib
from .device import DeviceArray, maximum_g, minimum_g, sum_g, mean_abs_g, max_abs_g
from . import kernels
from .solver import (DyrelParams, HydroThermal3D, HydroThermal3DBatch, dyrel_solve, reference_params,
                     production_params, SolveError)
from . import inputs

__all__ = ["DeviceArray", "kernels", "inputs", "DyrelParams", "HydroThermal3D", "HydroThermal3DBatch", "dyrel_solve",
           "reference_params",
           "production_params", "SolveError", "maximum_g", "minimum_g", "sum_g", "mean_abs_g", "max_abs_g"]
