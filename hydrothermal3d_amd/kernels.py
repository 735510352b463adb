"""Reference-named kernels on device arrays (the drop-in layer).

Each function here is one `@parallel Kernel!(args...)` launch of
tduretz/HydroThermal3D, with the same name (without "!") and the same argument
order, executed by the matching HIP kernel of libht3d.so.  Grid extents are
derived from the array shapes exactly as ParallelStencil does, and every shape
is checked against the reference's staggering before the launch (a mismatch
raises ValueError instead of reading out of bounds).

Staggering (cell grid nx x ny x nz): c = (nx,ny,nz), E = (nx+2,ny+2,nz+2),
X6 = (nx+6,ny+6,nz+6), V = (nx+1,ny+1,nz+1), Fx/Fy/Fz = one more along x/y/z.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .device import DeviceArray

_OFF = {
    "c": (0, 0, 0), "E": (2, 2, 2), "X6": (6, 6, 6), "V": (1, 1, 1),
    "Fx": (1, 0, 0), "Fy": (0, 1, 0), "Fz": (0, 0, 1),
}


def _grid_from(a: DeviceArray, kind: str) -> tuple[int, int, int]:
    o = _OFF[kind]
    return tuple(s - d for s, d in zip(a.shape, o))


def _check(a, kind: str, grid, name: str, argname: str, itemsize: int = 8) -> None:
    if not isinstance(a, DeviceArray):
        raise TypeError(f"{name}: argument {argname} must be a DeviceArray, got {type(a).__name__}")
    want = tuple(g + d for g, d in zip(grid, _OFF[kind]))
    if a.shape != want:
        raise ValueError(f"{name}: {argname} has shape {a.shape}, expected {want} ({kind} of grid {grid})")
    if a.dtype.itemsize != itemsize:
        raise ValueError(f"{name}: {argname} must be float{8 * itemsize}")


def _make(name: str, cname: str, kinds: list[tuple[str, str]], ref: str, fp32: bool = False):
    """kinds: [(argname, kind)] with kind in _OFF, 'd' (double) or 'i' (int).
    fp32: float32 arrays and float scalars (the _f32 entries)."""
    first = next(i for i, (_, k) in enumerate(kinds) if k in _OFF)

    def fn(*args):
        if len(args) != len(kinds):
            raise TypeError(f"{name} takes {len(kinds)} arguments ({', '.join(a for a, _ in kinds)}), got {len(args)}")
        grid = _grid_from(args[first], kinds[first][1])
        cargs = []
        for (argname, kind), v in zip(kinds, args):
            if kind in _OFF:
                _check(v, kind, grid, name, argname, 4 if fp32 else 8)
                cargs.append(C.c_void_p(v.ptr))
            elif kind == "W":   # (ncx+2, ncz+2) table of surface weights
                if not isinstance(v, DeviceArray) or v.shape != (grid[0] + 2, grid[2] + 2) or v.dtype.itemsize != 8:
                    raise ValueError(f"{name}: {argname} must be a float64 DeviceArray of shape "
                                     f"{(grid[0] + 2, grid[2] + 2)}")
                cargs.append(C.c_void_p(v.ptr))
            elif kind == "d":
                cargs.append(C.c_float(float(v)) if fp32 else C.c_double(float(v)))
            else:
                cargs.append(C.c_int(int(v)))
        _lib.call(cname, *cargs, *grid)

    fn.__name__ = name
    fn.__qualname__ = name
    fn.__doc__ = f"{name}!({', '.join(a for a, _ in kinds)}) -- {ref}"
    return fn


def _same_shape(name: str, cname: str, ref: str):
    def fn(A: DeviceArray, B: DeviceArray):
        if A.shape != B.shape:
            raise ValueError(f"{name}: shapes differ {A.shape} vs {B.shape}")
        _lib.call(cname, C.c_void_p(A.ptr), C.c_void_p(B.ptr), *A.shape)

    fn.__name__ = name
    fn.__doc__ = f"{name}!(A, B) -- {ref}"
    return fn


# --------------------------------------------------------------- tools/Macros.jl
ResetA = _same_shape("ResetA", "ht3d_ResetA", "tools/Macros.jl:171-177")
ArrayEqualArray = _same_shape("ArrayEqualArray", "ht3d_ArrayEqualArray", "tools/Weno5_Routines.jl:271-276")
Cpy_inn_to_all = _make("Cpy_inn_to_all", "ht3d_Cpy_inn_to_all", [("A", "c"), ("B", "E")], "tools/Macros.jl:187-192")
Cpy_all_to_inn = _make("Cpy_all_to_inn", "ht3d_Cpy_all_to_inn", [("A", "E"), ("B", "c")], "tools/Macros.jl:194-199")
Multiply = _make("Multiply", "ht3d_Multiply", [("a", "c"), ("b", "c"), ("c", "c")],
                 "Dimensional/kernels_HT3D_dimensional.jl:106-109")

# --------------------------------------------------------------- coefficients
InitThermal = _make("InitThermal", "ht3d_InitThermal",
                    [("Rt", "c"), ("kx", "Fx"), ("ky", "Fy"), ("kz", "Fz"), ("Tc_ex", "E"), ("ktv", "V"), ("_dt", "d")],
                    "Dimensionless/kernels_HT3D.jl:99-107")
InitConductivity = _make("InitConductivity", "ht3d_InitConductivity",
                         [("kx", "Fx"), ("ky", "Fy"), ("kz", "Fz"), ("ktv", "V")],
                         "Dimensional/kernels_HT3D_dimensional.jl:133-138")
InitDarcy = _make("InitDarcy", "ht3d_InitDarcy",
                  [("Ty", "Fy"), ("kx", "Fx"), ("ky", "Fy"), ("kz", "Fz"), ("Tc_ex", "E"), ("kfv", "V"), ("_dt", "d")],
                  "Dimensionless/kernels_HT3D.jl:82-90")
KillNorthSouthFluxes = _make("KillNorthSouthFluxes", "ht3d_KillNorthSouthFluxes", [("Ty", "Fy")],
                             "Dimensionless/HydroThermal3D_Alban_v6.jl:153-159")
ApplyViscosity = _make("ApplyViscosity", "ht3d_ApplyViscosity",
                       [("kx", "Fx"), ("ky", "Fy"), ("kz", "Fz"), ("Tc_ex", "E"), ("a", "d")],
                       "mu = exp(-a T) face mobility (no dimensionless counterpart; cf. kernels_HT3D_dimensional.jl)")
Compute_Rp = _make("Compute_Rp", "ht3d_Compute_Rp", [("Rp", "c"), ("ky", "Fy"), ("Ty", "Fy"), ("Ra", "d"), ("_dy", "d")],
                   "Dimensionless/kernels_HT3D.jl:92-97")

# --------------------------------------------------------------- BCs
SetTemperatureBCs = _make("SetTemperatureBCs", "ht3d_SetTemperatureBCs", [("Tc_ex", "E"), ("Tbot", "d"), ("Ttop", "d")],
                          "Dimensionless/HydroThermal3D_Alban_v6.jl:108-118")
SetPressureBCs = _make("SetPressureBCs", "ht3d_SetPressureBCs", [("Pc_ex", "E"), ("Pbot", "d"), ("Ptop", "d")],
                       "Dimensionless/HydroThermal3D_Alban_v6.jl:120-130")
SetPressureBCs_v7 = _make("SetPressureBCs_v7", "ht3d_SetPressureBCs_v7", [("Pc_ex", "E"), ("Pbot", "d"), ("Ptop", "d")],
                          "Dimensionless/kernels_HT3D.jl:144-154")

# --------------------------------------------------------------- fluxes / residuals / updates
ComputeFlux = _make("ComputeFlux", "ht3d_ComputeFlux",
                    [("qx", "Fx"), ("qy", "Fy"), ("qz", "Fz"), ("kx", "Fx"), ("ky", "Fy"), ("kz", "Fz"), ("A", "E"),
                     ("_dx", "d"), ("_dy", "d"), ("_dz", "d")], "Dimensionless/kernels_HT3D.jl:109-117")
UpdateT = _make("UpdateT", "ht3d_UpdateT",
                [("F", "c"), ("T", "E"), ("R", "c"), ("qx", "Fx"), ("qy", "Fy"), ("qz", "Fz"), ("_dt", "d"),
                 ("dtau", "d"), ("_dx", "d"), ("_dy", "d"), ("_dz", "d")], "Dimensionless/kernels_HT3D.jl:119-126")
ResidualDiffusion = _make("ResidualDiffusion", "ht3d_ResidualDiffusion",
                          [("F", "c"), ("R", "c"), ("qx", "Fx"), ("qy", "Fy"), ("qz", "Fz"), ("_dx", "d"), ("_dy", "d"),
                           ("_dz", "d")], "Dimensionless/kernels_HT3D.jl:128-134")
ResidualDiffusionTransient = _make(
    "ResidualDiffusionTransient", "ht3d_ResidualDiffusionTransient",
    [("F", "c"), ("T", "E"), ("R", "c"), ("qx", "Fx"), ("qy", "Fy"), ("qz", "Fz"), ("_dt", "d"), ("_dx", "d"),
     ("_dy", "d"), ("_dz", "d")], "Dimensionless/HydroThermal3D_Alban_v6.jl:84-90")
ResidualDiffusionPC = _make(
    "ResidualDiffusionPC", "ht3d_ResidualDiffusionPC",
    [("F", "c"), ("R", "c"), ("qx", "Fx"), ("qy", "Fy"), ("qz", "Fz"), ("PC", "c"), ("_dx", "d"), ("_dy", "d"),
     ("_dz", "d")], "kernels_HT3D.jl:128-134 then F /= PC (kernels_HT3D_dimensional.jl:283)")
ResidualDiffusionTransientPC = _make(
    "ResidualDiffusionTransientPC", "ht3d_ResidualDiffusionTransientPC",
    [("F", "c"), ("T", "E"), ("R", "c"), ("qx", "Fx"), ("qy", "Fy"), ("qz", "Fz"), ("PC", "c"), ("_dt", "d"),
     ("_dx", "d"), ("_dy", "d"), ("_dz", "d")], "Alban_v6:84-90 then F /= PC (kernels_HT3D_dimensional.jl:175)")
UpdateP = _make("UpdateP", "ht3d_UpdateP", [("F0", "c"), ("P", "E"), ("F", "c"), ("dampx", "d"), ("dtau", "d")],
                "Dimensionless/kernels_HT3D.jl:136-142")
DampedUpdate = _make("DampedUpdate", "ht3d_DampedUpdate", [("F0", "c"), ("X", "E"), ("F", "c"), ("dampx", "d"), ("_dtau", "d")],
                     "Dimensional/kernels_HT3D_dimensional.jl:316-320")
DYRELUpdate = _make("DYRELUpdate", "ht3d_DYRELUpdate",
                    [("F0", "c"), ("X", "E"), ("F", "c"), ("h1", "d"), ("h2", "d"), ("dtau", "d")],
                    "Dimensional/kernels_HT3D_dimensional.jl:322-326")
SwapDYREL = _make("SwapDYREL", "ht3d_SwapDYREL", [("Xc", "c"), ("Yc", "c"), ("X0c_ex", "E"), ("Y0c", "c")],
                  "Dimensional/kernels_HT3D_dimensional.jl:122-126")
GershgorinPoisson = _make(
    "GershgorinPoisson", "ht3d_GershgorinPoisson",
    [("G", "c"), ("D", "c"), ("rhoC", "c"), ("kx", "Fx"), ("ky", "Fy"), ("kz", "Fz"), ("transient", "d"), ("dt", "d"),
     ("dx", "d"), ("dy", "d"), ("dz", "d")], "Dimensional/kernels_HT3D_dimensional.jl:207-221")
GershgorinHydro = _make(
    "GershgorinHydro", "ht3d_GershgorinHydro",
    [("G", "c"), ("phidrhodP", "c"), ("kx", "Fx"), ("ky", "Fy"), ("kz", "Fz"), ("transient", "d"), ("dt", "d"),
     ("dx", "d"), ("dy", "d"), ("dz", "d")], "Dimensional/kernels_HT3D_dimensional.jl:335-347")
RayleighQuotientNumerator = _make(
    "RayleighQuotientNumerator", "ht3d_RayleighQuotientNumerator",
    [("G", "c"), ("Fc0", "c"), ("Fcit", "c"), ("Fc", "c"), ("dtau", "d")],
    "Dimensional/kernels_HT3D_dimensional.jl:349-355")
RayleighQuotientDenominator = _make(
    "RayleighQuotientDenominator", "ht3d_RayleighQuotientDenominator", [("G", "c"), ("Fc0", "c"), ("dtau", "d")],
    "Dimensional/kernels_HT3D_dimensional.jl:357-363")
ResidualTemperatureLinearised = _make(
    "ResidualTemperatureLinearised", "ht3d_ResidualTemperatureLinearised",
    [("F", "c"), ("Tc_ex", "E"), ("Tc_old", "c"), ("rhoC_eff", "c"), ("phc_ex", "E"), ("PC", "c"), ("qx", "Fx"),
     ("qy", "Fy"), ("qz", "Fz"), ("Qt", "d"), ("transient", "d"), ("_dt", "d"), ("_dx", "d"), ("_dy", "d"),
     ("_dz", "d")], "Dimensional/kernels_HT3D_dimensional.jl:162-181")
ResidualFluidPressure = _make(
    "ResidualFluidPressure", "ht3d_ResidualFluidPressure",
    [("F", "c"), ("phc_ex", "E"), ("phirhof", "c"), ("phirhof0", "c"), ("PC", "c"), ("qx", "Fx"), ("qy", "Fy"),
     ("qz", "Fz"), ("transient", "d"), ("_dt", "d"), ("_dx", "d"), ("_dy", "d"), ("_dz", "d")],
    "Dimensional/kernels_HT3D_dimensional.jl:274-289")
ComputeDarcyFlux = _make(
    "ComputeDarcyFlux", "ht3d_ComputeDarcyFlux",
    [("qx", "Fx"), ("qy", "Fy"), ("qz", "Fz"), ("rhofv", "V"), ("kx", "Fx"), ("ky", "Fy"), ("kz", "Fz"), ("Pf", "E"),
     ("g", "d"), ("_dx", "d"), ("_dy", "d"), ("_dz", "d")], "Dimensional/kernels_HT3D_dimensional.jl:266-272")
Init_vel = _make(
    "Init_vel", "ht3d_Init_vel",
    [("Vx", "Fx"), ("Vy", "Fy"), ("Vz", "Fz"), ("kx", "Fx"), ("ky", "Fy"), ("kz", "Fz"), ("Pc_ex", "E"), ("Ty", "Fy"),
     ("Ra", "d"), ("_dx", "d"), ("_dy", "d"), ("_dz", "d")], "Dimensionless/kernels_HT3D.jl:73-80")


# --------------------------------------------------------------- dimensional (v18) property kernels
_D = "Dimensional/kernels_HT3D_dimensional.jl"
ComputeThermalConductivity = _make(
    "ComputeThermalConductivity", "ht3d_ComputeThermalConductivity",
    [("ktv", "V"), ("Tce", "E"), ("phv", "V"), ("phi0", "d"), ("scale_kt", "d"), ("scale_T", "d")], _D + ":40-49")
ComputeHydroConductivity = _make(
    "ComputeHydroConductivity", "ht3d_ComputeHydroConductivity",
    [("kfv", "V"), ("Tce", "E"), ("Pce", "E"), ("phv", "V"), ("ymin", "d"), ("dy", "d"), ("k_fact", "d"),
     ("delta", "d"), ("scale_sigma", "d"), ("scale_T", "d"), ("scale_L", "d"), ("scale_t", "d"), ("scale_eta", "d"),
     ("mode", "i")], _D + ":51-83")
ComputeFluidDensity = _make(
    "ComputeFluidDensity", "ht3d_ComputeFluidDensity",
    [("rhofv", "V"), ("Tce", "E"), ("Pce", "E"), ("phv", "V"), ("scale_sigma", "d"), ("scale_T", "d"),
     ("scale_rho", "d")], _D + ":85-99")
SmoothConductivityV2C = _make("SmoothConductivityV2C", "ht3d_SmoothConductivityV2C", [("ktc", "c"), ("ktv", "V")],
                              _D + ":110-113")
SmoothConductivityC2V = _make("SmoothConductivityC2V", "ht3d_SmoothConductivityC2V", [("ktv", "V"), ("ktc", "c")],
                              _D + ":115-118")
ComputeRhoCeffective = _make(
    "ComputeRhoCeffective", "ht3d_ComputeRhoCeffective",
    [("rhoC_eff", "c"), ("Tc_ex", "E"), ("Pc_ex", "E"), ("phc_ex", "E"), ("phi_i", "d"), ("sca_T", "d"),
     ("sca_sigma", "d"), ("sca_C", "d"), ("sca_rho", "d")], _D + ":148-160 (ComputeρCeffective!)")
_PHI = [("phc_ex", "E"), ("phi_i", "d"), ("scale_rho", "d"), ("scale_T", "d"), ("scale_sigma", "d")]
PhiRhof = _make("PhiRhof", "ht3d_PhiRhof", [("phirhofc", "c"), ("Tc_ex", "E"), ("Pc_ex", "E")] + _PHI,
                _D + ":225-235 (ϕρf)")
PhiRhof_old = _make("PhiRhof_old", "ht3d_PhiRhof_old", [("phirhofc", "c"), ("Tc0", "c"), ("Pc0", "c")] + _PHI,
                    _D + ":237-247 (ϕρf_old)")
PhiDrhoDP = _make("PhiDrhoDP", "ht3d_PhiDrhoDP", [("phidrhodP", "c"), ("Tc_ex", "E"), ("Pc_ex", "E")] + _PHI,
                  _D + ":249-259 (ϕdρdP!)")
dim_Init_vel = _make(
    "dim_Init_vel", "ht3d_dim_Init_vel",
    [("Vx", "Fx"), ("Vy", "Fy"), ("Vz", "Fz"), ("qx", "Fx"), ("qy", "Fy"), ("qz", "Fz"), ("rhofv", "V")],
    _D + ":328-332 (the dimensional Init_vel!)")
dim_SetTemperatureBCs = _make(
    "dim_SetTemperatureBCs", "ht3d_dim_SetTemperatureBCs",
    [("Tc_ex", "E"), ("ky", "Fy"), ("wN", "W"), ("phc_ex", "E")] +
    [(n, "d") for n in ("qyS", "_dy", "kS", "TS", "TN", "y_plateau", "a2", "b2", "dTdy", "dx", "dy", "sticky_air",
                        "sca_T", "sca_kT", "phi_i")], _D + ":183-205 (the dimensional SetTemperatureBCs!)")
dim_SetPressureBCs = _make(
    "dim_SetPressureBCs", "ht3d_dim_SetPressureBCs",
    [("Pc_ex", "E"), ("wN", "W"), ("phc_ex", "E")] +
    [(n, "d") for n in ("PS", "PN", "y_plateau", "a2", "b2", "rhof", "g", "dx", "dy", "sticky_air")],
    _D + ":291-314 (the dimensional SetPressureBCs!)")


# --------------------------------------------------------------- output fields (kernels_HT3D_postprocessing.jl)
_PP = "Dimensional/kernels_HT3D_postprocessing.jl"
Phase = _make("Phase", "ht3d_Phase", [("X", "c"), ("phc_ex", "E")], _PP + ":1-10")
Pressure = _make("Pressure", "ht3d_Pressure", [("X", "c"), ("Pc_ex", "E"), ("phc_ex", "E"), ("scale_sigma", "d")],
                 _PP + ":12-21")
Temperature = _make("Temperature", "ht3d_Temperature", [("X", "c"), ("Tc_ex", "E"), ("phc_ex", "E"), ("scale_T", "d")],
                    _PP + ":23-32")
Velocity = _make("Velocity", "ht3d_Velocity",
                 [("X", "c"), ("Y", "c"), ("Z", "c"), ("Vx", "Fx"), ("Vy", "Fy"), ("Vz", "Fz"), ("phc_ex", "E"),
                  ("scale_V", "d")], _PP + ":34-45")
EffectiveThermalConductivity = _make(
    "EffectiveThermalConductivity", "ht3d_EffectiveThermalConductivity",
    [("X", "c"), ("Tc_ex", "E"), ("phi0", "d"), ("phc_ex", "E"), ("scale_T", "d")], _PP + ":47-57")
EffectiveHeatCapacity = _make(
    "EffectiveHeatCapacity", "ht3d_EffectiveHeatCapacity",
    [("Ceff", "c"), ("Tc_ex", "E"), ("Pc_ex", "E"), ("phc_ex", "E"), ("phi_i", "d"), ("sca_T", "d"),
     ("sca_sigma", "d"), ("sca_C", "d"), ("sca_rho", "d")], _PP + ":60-70")
Permeability = _make("Permeability", "ht3d_Permeability",
                     [("X", "c"), ("k_fact", "d"), ("ymin", "d"), ("dy", "d"), ("phc_ex", "E"), ("delta", "d"),
                      ("scale_L", "d")], _PP + ":72-85")
FluidDensity = _make("FluidDensity", "ht3d_FluidDensity",
                     [("X", "c"), ("Tc_ex", "E"), ("Pc_ex", "E"), ("phc_ex", "E"), ("scale_T", "d"),
                      ("scale_sigma", "d")], _PP + ":87-98")
Viscosity = _make("Viscosity", "ht3d_Viscosity", [("X", "c"), ("Tc_ex", "E"), ("phc_ex", "E"), ("scale_T", "d")],
                  _PP + ":100-109")
Peclet = _make("Peclet", "ht3d_Peclet",
               [("X", "c"), ("Tc_ex", "E"), ("Pc_ex", "E"), ("Vx", "Fx"), ("Vy", "Fy"), ("Vz", "Fz"),
                ("phc_ex", "E"), ("phi0", "d"), ("scale_T", "d"), ("scale_sigma", "d"), ("scale_V", "d"),
                ("scale_L", "d")], _PP + ":111-128")

# --------------------------------------------------------------- WENO5 (tools/Weno5_Routines*.jl)
VxPlusMinus = _make("VxPlusMinus", "ht3d_VxPlusMinus", [("Vxm", "c"), ("Vxp", "c"), ("Vx", "Fx")],
                    "tools/Weno5_Routines.jl:5-16")
VyPlusMinus = _make("VyPlusMinus", "ht3d_VyPlusMinus", [("Vym", "c"), ("Vyp", "c"), ("Vy", "Fy")],
                    "tools/Weno5_Routines.jl:18-29")
VzPlusMinus = _make("VzPlusMinus", "ht3d_VzPlusMinus", [("Vzm", "c"), ("Vzp", "c"), ("Vz", "Fz")],
                    "tools/Weno5_Routines.jl:31-42")
dFdx_Weno5 = _make("dFdx_Weno5", "ht3d_dFdx_Weno5",
                   [("dFdxi", "c"), ("V1", "c"), ("V2", "c"), ("V3", "c"), ("V4", "c"), ("V5", "c")],
                   "tools/Weno5_Routines.jl:44-68")
_GRAD = [("v1", "c"), ("v2", "c"), ("v3", "c"), ("v4", "c"), ("v5", "c"), ("Fc_exxx", "X6"), ("_dx", "d"), ("_dy", "d"),
         ("_dz", "d")]
Gradients_minus_x_Weno5 = _make("Gradients_minus_x_Weno5", "ht3d_Gradients_minus_x_Weno5", _GRAD,
                                "tools/Weno5_Routines.jl:70-78")
Gradients_plus_x_Weno5 = _make("Gradients_plus_x_Weno5", "ht3d_Gradients_plus_x_Weno5", _GRAD,
                               "tools/Weno5_Routines.jl:80-88")
Gradients_minus_y_Weno5 = _make("Gradients_minus_y_Weno5", "ht3d_Gradients_minus_y_Weno5", _GRAD,
                                "tools/Weno5_Routines.jl:90-98")
Gradients_plus_y_Weno5 = _make("Gradients_plus_y_Weno5", "ht3d_Gradients_plus_y_Weno5", _GRAD,
                               "tools/Weno5_Routines.jl:100-108")
Gradients_minus_z_Weno5 = _make("Gradients_minus_z_Weno5", "ht3d_Gradients_minus_z_Weno5", _GRAD,
                                "tools/Weno5_Routines.jl:110-118")
Gradients_plus_z_Weno5 = _make("Gradients_plus_z_Weno5", "ht3d_Gradients_plus_z_Weno5", _GRAD,
                               "tools/Weno5_Routines.jl:120-128")


def _boundaries(axis: str, name: str, ref: str):
    def fn(Fc_exxx: DeviceArray, Fc: DeviceArray, type_lo: int, val_lo: float, type_hi: int, val_hi: float):
        grid = _grid_from(Fc_exxx, "X6")
        _check(Fc_exxx, "X6", grid, name, "Fc_exxx")
        if Fc.shape == tuple(g + 2 for g in grid):
            entry = f"ht3d_Boundaries_{axis}_Weno5_v2"    # tools/Weno5_Routines_v2.jl (Fc is the extended array)
        else:
            _check(Fc, "c", grid, name, "Fc")
            entry = f"ht3d_Boundaries_{axis}_Weno5"       # tools/Weno5_Routines.jl (Fc is the cell array)
        _lib.call(entry, C.c_void_p(Fc_exxx.ptr), C.c_void_p(Fc.ptr), int(type_lo), C.c_double(val_lo),
                  int(type_hi), C.c_double(val_hi), *grid)

    fn.__name__ = name
    fn.__doc__ = f"{name}!(Fc_exxx, Fc, type_lo, val_lo, type_hi, val_hi) -- {ref}"
    return fn


Boundaries_x_Weno5 = _boundaries("x", "Boundaries_x_Weno5", "tools/Weno5_Routines.jl:130-170, _v2.jl:221-263")
Boundaries_y_Weno5 = _boundaries("y", "Boundaries_y_Weno5", "tools/Weno5_Routines.jl:172-212, _v2.jl:265-305")
Boundaries_z_Weno5 = _boundaries("z", "Boundaries_z_Weno5", "tools/Weno5_Routines.jl:214-255, _v2.jl:307-347")

_Advect_v1 = _make("Advect", "ht3d_Advect",
                   [("Fc", "c"), ("Vxp", "c"), ("dTdxm", "c"), ("Vxm", "c"), ("dTdxp", "c"), ("dt", "d")],
                   "tools/Weno5_Routines.jl:257-262")


def Advect(*args):
    """Advect!: the v1 stage update (6 arguments, tools/Weno5_Routines.jl:257-262) or
    the fused v2 kernel (8 arguments: Fc, X, V, axis, dt, _dx, _dy, _dz;
    tools/Weno5_Routines_v2.jl:55-212), dispatched on arity like Julia's methods."""
    if len(args) == 6:
        return _Advect_v1(*args)
    if len(args) != 8:
        raise TypeError("Advect takes 6 (v1) or 8 (v2) arguments")
    Fc, X, V, axis, dt, _dx, _dy, _dz = args
    grid = _grid_from(Fc, "E")
    _check(Fc, "E", grid, "Advect", "Fc")
    _check(X, "X6", grid, "Advect", "X")
    kind = {1: "Fx", 2: "Fy", 3: "Fz"}[int(axis)]
    _check(V, kind, grid, "Advect", "V")
    _lib.call("ht3d_Advect_v2", C.c_void_p(Fc.ptr), C.c_void_p(X.ptr), C.c_void_p(V.ptr), int(axis), C.c_double(dt),
              C.c_double(_dx), C.c_double(_dy), C.c_double(_dz), *grid)


def TimeAveraging(Fc: DeviceArray, Fold: DeviceArray, order: float):
    """TimeAveraging!: on the cell array (tools/Weno5_Routines.jl:264-269) or on the
    interior of the extended array (tools/Weno5_Routines_v2.jl:214-219)."""
    grid = _grid_from(Fold, "c")
    _check(Fold, "c", grid, "TimeAveraging", "Fold")
    if Fc.shape == Fold.shape:
        _lib.call("ht3d_TimeAveraging", C.c_void_p(Fc.ptr), C.c_void_p(Fold.ptr), C.c_double(order), *grid)
    else:
        _check(Fc, "E", grid, "TimeAveraging", "Fc")
        _lib.call("ht3d_TimeAveraging_v2", C.c_void_p(Fc.ptr), C.c_void_p(Fold.ptr), C.c_double(order), *grid)


def AdvectWithWeno5_v2(Tc_ex, Tc_exxx, Told, Vx, Vy, Vz, dt, _dx, _dy, _dz, Ttop, Tbot):
    """AdvectWithWeno5_v2 (tools/Weno5_Routines_v2.jl:1-50) as one fused call.
    Tc_exxx and Told are accepted for signature compatibility; the fused passes
    need neither (ghosts are mirrored inside the stencil gather)."""
    grid = _grid_from(Tc_ex, "E")
    for nm, a, k in (("Vx", Vx, "Fx"), ("Vy", Vy, "Fy"), ("Vz", Vz, "Fz")):
        _check(a, k, grid, "AdvectWithWeno5_v2", nm)
    _lib.call("ht3d_weno5_advect", C.c_void_p(Tc_ex.ptr), C.c_void_p(Vx.ptr), C.c_void_p(Vy.ptr), C.c_void_p(Vz.ptr),
              C.c_double(dt), C.c_double(_dx), C.c_double(_dy), C.c_double(_dz), C.c_double(Ttop), C.c_double(Tbot),
              2, *grid)


def AdvectWithWeno5(Tc_ex, Vx, Vy, Vz, dt, _dx, _dy, _dz, Ttop, Tbot):
    """AdvectWithWeno5 (Dimensionless/kernels_HT3D.jl:156-233, v1 stencils) after
    Init_vel!, as one fused call; the fifteen scratch arrays of the reference
    signature are not needed."""
    grid = _grid_from(Tc_ex, "E")
    for nm, a, k in (("Vx", Vx, "Fx"), ("Vy", Vy, "Fy"), ("Vz", Vz, "Fz")):
        _check(a, k, grid, "AdvectWithWeno5", nm)
    _lib.call("ht3d_weno5_advect", C.c_void_p(Tc_ex.ptr), C.c_void_p(Vx.ptr), C.c_void_p(Vy.ptr), C.c_void_p(Vz.ptr),
              C.c_double(dt), C.c_double(_dx), C.c_double(_dy), C.c_double(_dz), C.c_double(Ttop), C.c_double(Tbot),
              1, *grid)


# ------------------------------------------------ FP32 inner-sweep kernels
_F32 = {
    "SetTemperatureBCs_f32": [("Tc_ex", "E"), ("Tbot", "d"), ("Ttop", "d")],
    "SetPressureBCs_f32": [("Pc_ex", "E"), ("Pbot", "d"), ("Ptop", "d")],
    "ComputeFlux_f32": [("qx", "Fx"), ("qy", "Fy"), ("qz", "Fz"), ("kx", "Fx"), ("ky", "Fy"), ("kz", "Fz"),
                        ("A", "E"), ("_dx", "d"), ("_dy", "d"), ("_dz", "d")],
    "ResidualDiffusion_f32": [("F", "c"), ("R", "c"), ("qx", "Fx"), ("qy", "Fy"), ("qz", "Fz"), ("_dx", "d"),
                              ("_dy", "d"), ("_dz", "d")],
    "ResidualDiffusionTransient_f32": [("F", "c"), ("T", "E"), ("R", "c"), ("qx", "Fx"), ("qy", "Fy"), ("qz", "Fz"),
                                       ("_dt", "d"), ("_dx", "d"), ("_dy", "d"), ("_dz", "d")],
    "ResidualDiffusionPC_f32": [("F", "c"), ("R", "c"), ("qx", "Fx"), ("qy", "Fy"), ("qz", "Fz"), ("PC", "c"),
                                ("_dx", "d"), ("_dy", "d"), ("_dz", "d")],
    "ResidualDiffusionTransientPC_f32": [("F", "c"), ("T", "E"), ("R", "c"), ("qx", "Fx"), ("qy", "Fy"),
                                         ("qz", "Fz"), ("PC", "c"), ("_dt", "d"), ("_dx", "d"), ("_dy", "d"),
                                         ("_dz", "d")],
    "DYRELUpdate_f32": [("F0", "c"), ("X", "E"), ("F", "c"), ("h1", "d"), ("h2", "d"), ("dtau", "d")],
    "GershgorinPoisson_f32": [("G", "c"), ("D", "c"), ("rhoC", "c"), ("kx", "Fx"), ("ky", "Fy"), ("kz", "Fz"),
                              ("transient", "d"), ("dt", "d"), ("dx", "d"), ("dy", "d"), ("dz", "d")],
}
for _n, _k in _F32.items():
    globals()[_n] = _make(_n, "ht3d_" + _n, _k, "FP32 inner sweep of the defect-correction solve", fp32=True)


def to_f32(out: DeviceArray, src: DeviceArray) -> None:
    """FP64 -> FP32 copy of a whole array (same shape)."""
    if out.shape != src.shape or out.dtype.itemsize != 4 or src.dtype.itemsize != 8:
        raise ValueError("to_f32: out must be float32 and src float64 of the same shape")
    _lib.call("ht3d_to_f32", C.c_void_p(out.ptr), C.c_void_p(src.ptr), C.c_size_t(int(np.prod(src.shape))))


def add_f32(X: DeviceArray, e: DeviceArray) -> None:
    """X += e on the interior of an FP64 E array (e: FP32 E array)."""
    grid = _grid_from(X, "E")
    _check(X, "E", grid, "add_f32", "X")
    _check(e, "E", grid, "add_f32", "e", 4)
    _lib.call("ht3d_add_f32", C.c_void_p(X.ptr), C.c_void_p(e.ptr), *grid)


def velocity_dt_nu(Vx, Vy, Vz, kx, ky, kz, Pc_ex, Ty, Ra, dx, dy, dz, cfl, Tc_ex=None, kty=None, Lx=1.0, Lz=1.0):
    """Init_vel! (kernels_HT3D.jl:73-80), max|V|, the advective time step
    cfl*min(h)/max|V| (HydroThermal3D_v18.jl:554) and, when Tc_ex and kty are
    given, Nu of HydroThermal3D_Alban_v6.jl:345-346, in one pass.  Returns
    (vmax, dt, Nu)."""
    grid = _grid_from(Pc_ex, "E")
    for nm, a, k in (("Vx", Vx, "Fx"), ("Vy", Vy, "Fy"), ("Vz", Vz, "Fz"), ("kx", kx, "Fx"), ("ky", ky, "Fy"),
                     ("kz", kz, "Fz"), ("Ty", Ty, "Fy")):
        _check(a, k, grid, "velocity_dt_nu", nm)
    if (Tc_ex is None) != (kty is None):
        raise ValueError("velocity_dt_nu: give both Tc_ex and kty, or neither")
    if Tc_ex is not None:
        _check(Tc_ex, "E", grid, "velocity_dt_nu", "Tc_ex")
        _check(kty, "Fy", grid, "velocity_dt_nu", "kty")
    out = (C.c_double * 3)()
    _lib.call("ht3d_velocity_dt_nu", C.c_void_p(Vx.ptr), C.c_void_p(Vy.ptr), C.c_void_p(Vz.ptr), C.c_void_p(kx.ptr),
              C.c_void_p(ky.ptr), C.c_void_p(kz.ptr), C.c_void_p(Pc_ex.ptr), C.c_void_p(Ty.ptr), C.c_double(Ra),
              C.c_double(dx), C.c_double(dy), C.c_double(dz), C.c_double(cfl),
              C.c_void_p(Tc_ex.ptr if Tc_ex is not None else None), C.c_void_p(kty.ptr if kty is not None else None),
              C.c_double(Lx), C.c_double(Lz), out, *grid)
    return out[0], out[1], out[2]


__all__ = [n for n in dir() if not n.startswith("_") and n[0].isupper()] + ["velocity_dt_nu", "to_f32", "add_f32"]
