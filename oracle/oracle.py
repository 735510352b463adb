"""Python binding of the CPU oracle (oracle/libht3d_oracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, to compare results with.  The product path never imports
this module.

Functions carry the reference kernel names (without "!") and argument order and
take Fortran-ordered float64 NumPy arrays with the reference's staggered shapes;
each forwards to the C restatement in ht3d_oracle.c.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
# HT3D_ORACLE_LIB selects another build of the same sources (the sanitized
# one, `make -C oracle sanitize`); default: the optimised build.
LIB = Path(os.environ["HT3D_ORACLE_LIB"]) if os.environ.get("HT3D_ORACLE_LIB") else HERE / "libht3d_oracle.so"

_OFF = {"c": (0, 0, 0), "E": (2, 2, 2), "X6": (6, 6, 6), "V": (1, 1, 1), "Fx": (1, 0, 0), "Fy": (0, 1, 0),
        "Fz": (0, 0, 1)}


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


_L = None


def lib() -> C.CDLL:
    global _L
    if _L is None:
        if not LIB.exists():
            build()
        _L = C.CDLL(str(LIB))
        _L.o_reduce.restype = C.c_double
        _L.o_reduce.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t, C.c_int]
        _L.o_weno5_js.restype = C.c_double
        _L.o_weno5_js.argtypes = [C.c_double] * 5
        _L.o_hrl_create.restype = C.c_void_p
        _L.o_hrl_create.argtypes = [C.c_int, C.c_int, C.c_int] + [C.c_double] * 5
        _L.o_hrl_field.restype = C.c_void_p
        _L.o_hrl_field.argtypes = [C.c_void_p, C.c_char_p]
        _L.o_hrl_scalar.restype = C.c_double
        _L.o_hrl_scalar.argtypes = [C.c_void_p, C.c_int]
        _L.o_hrl_step.argtypes = [C.c_void_p]
        _L.o_hrl_destroy.argtypes = [C.c_void_p]
        _L.o_hrl_set_params.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
        _L.o_hrl_get_reports.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        _L.o_hrl_set_dt.argtypes = [C.c_void_p, C.c_double]
        _L.o_hrl_set_viscosity.argtypes = [C.c_void_p, C.c_double]
        _L.o_hrl_set_dt_rule.argtypes = [C.c_void_p, C.c_int]
        _L.o_hrl_n_adv.argtypes = [C.c_void_p]
        _L.o_set_threads.argtypes = [C.c_int]
        _L.o_get_threads.restype = C.c_int
        _L.o_dyrel_solve.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        _L.o_eos.restype = C.c_double
        _L.o_eos.argtypes = [C.c_int, C.c_double, C.c_double]
    return _L


def _ptr(a: np.ndarray) -> C.c_void_p:
    if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.f_contiguous):
        raise TypeError("oracle arrays must be Fortran-ordered float64 numpy arrays")
    return C.c_void_p(a.ctypes.data)


def _grid(a: np.ndarray, kind: str):
    return tuple(s - d for s, d in zip(a.shape, _OFF[kind]))


def _make(name: str, cname: str, kinds):
    first = next(i for i, (_, k) in enumerate(kinds) if k in _OFF)

    def fn(*args):
        assert len(args) == len(kinds), f"{name}: expected {len(kinds)} args"
        grid = _grid(args[first], kinds[first][1])
        cargs = []
        for (argname, kind), v in zip(kinds, args):
            if kind == "W":
                want = (grid[0] + 2, grid[2] + 2)
                if v.shape != want:
                    raise ValueError(f"{name}: {argname} shape {v.shape} != {want}")
                cargs.append(_ptr(v))
            elif kind in _OFF:
                want = tuple(g + d for g, d in zip(grid, _OFF[kind]))
                if v.shape != want:
                    raise ValueError(f"{name}: {argname} shape {v.shape} != {want}")
                cargs.append(_ptr(v))
            elif kind == "d":
                cargs.append(C.c_double(float(v)))
            else:
                cargs.append(C.c_int(int(v)))
        getattr(lib(), cname)(*cargs, *grid)

    fn.__name__ = name
    return fn


def ResetA(A, B):
    lib().o_ResetA(_ptr(A), _ptr(B), *A.shape)


def ArrayEqualArray(A, B):
    lib().o_ArrayEqualArray(_ptr(A), _ptr(B), *A.shape)


Cpy_inn_to_all = _make("Cpy_inn_to_all", "o_Cpy_inn_to_all", [("A", "c"), ("B", "E")])
Cpy_all_to_inn = _make("Cpy_all_to_inn", "o_Cpy_all_to_inn", [("A", "E"), ("B", "c")])
Multiply = _make("Multiply", "o_Multiply", [("a", "c"), ("b", "c"), ("c", "c")])
InitThermal = _make("InitThermal", "o_InitThermal",
                    [("Rt", "c"), ("kx", "Fx"), ("ky", "Fy"), ("kz", "Fz"), ("Tc_ex", "E"), ("ktv", "V"), ("_dt", "d")])
InitConductivity = _make("InitConductivity", "o_InitConductivity", [("kx", "Fx"), ("ky", "Fy"), ("kz", "Fz"), ("ktv", "V")])
InitDarcy = _make("InitDarcy", "o_InitDarcy",
                  [("Ty", "Fy"), ("kx", "Fx"), ("ky", "Fy"), ("kz", "Fz"), ("Tc_ex", "E"), ("kfv", "V"), ("_dt", "d")])
KillNorthSouthFluxes = _make("KillNorthSouthFluxes", "o_KillNorthSouthFluxes", [("Ty", "Fy")])
ApplyViscosity = _make("ApplyViscosity", "o_ApplyViscosity",
                       [("kx", "Fx"), ("ky", "Fy"), ("kz", "Fz"), ("Tc_ex", "E"), ("a", "d")])
Compute_Rp = _make("Compute_Rp", "o_Compute_Rp", [("Rp", "c"), ("ky", "Fy"), ("Ty", "Fy"), ("Ra", "d"), ("_dy", "d")])
SetTemperatureBCs = _make("SetTemperatureBCs", "o_SetTemperatureBCs", [("Tc_ex", "E"), ("Tbot", "d"), ("Ttop", "d")])
SetPressureBCs = _make("SetPressureBCs", "o_SetPressureBCs", [("Pc_ex", "E"), ("Pbot", "d"), ("Ptop", "d")])
SetPressureBCs_v7 = _make("SetPressureBCs_v7", "o_SetPressureBCs_v7", [("Pc_ex", "E"), ("Pbot", "d"), ("Ptop", "d")])
ComputeFlux = _make("ComputeFlux", "o_ComputeFlux",
                    [("qx", "Fx"), ("qy", "Fy"), ("qz", "Fz"), ("kx", "Fx"), ("ky", "Fy"), ("kz", "Fz"), ("A", "E"),
                     ("_dx", "d"), ("_dy", "d"), ("_dz", "d")])
UpdateT = _make("UpdateT", "o_UpdateT", [("F", "c"), ("T", "E"), ("R", "c"), ("qx", "Fx"), ("qy", "Fy"), ("qz", "Fz"),
                                         ("_dt", "d"), ("dtau", "d"), ("_dx", "d"), ("_dy", "d"), ("_dz", "d")])
ResidualDiffusion = _make("ResidualDiffusion", "o_ResidualDiffusion",
                          [("F", "c"), ("R", "c"), ("qx", "Fx"), ("qy", "Fy"), ("qz", "Fz"), ("_dx", "d"), ("_dy", "d"),
                           ("_dz", "d")])
ResidualDiffusionTransient = _make("ResidualDiffusionTransient", "o_ResidualDiffusionTransient",
                                   [("F", "c"), ("T", "E"), ("R", "c"), ("qx", "Fx"), ("qy", "Fy"), ("qz", "Fz"),
                                    ("_dt", "d"), ("_dx", "d"), ("_dy", "d"), ("_dz", "d")])
ResidualDiffusionPC = _make("ResidualDiffusionPC", "o_ResidualDiffusionPC",
                            [("F", "c"), ("R", "c"), ("qx", "Fx"), ("qy", "Fy"), ("qz", "Fz"), ("PC", "c"), ("_dx", "d"),
                             ("_dy", "d"), ("_dz", "d")])
ResidualDiffusionTransientPC = _make("ResidualDiffusionTransientPC", "o_ResidualDiffusionTransientPC",
                                     [("F", "c"), ("T", "E"), ("R", "c"), ("qx", "Fx"), ("qy", "Fy"), ("qz", "Fz"),
                                      ("PC", "c"), ("_dt", "d"), ("_dx", "d"), ("_dy", "d"), ("_dz", "d")])
UpdateP = _make("UpdateP", "o_UpdateP", [("F0", "c"), ("P", "E"), ("F", "c"), ("dampx", "d"), ("dtau", "d")])
DampedUpdate = UpdateP
DYRELUpdate = _make("DYRELUpdate", "o_DYRELUpdate",
                    [("F0", "c"), ("X", "E"), ("F", "c"), ("h1", "d"), ("h2", "d"), ("dtau", "d")])
SwapDYREL = _make("SwapDYREL", "o_SwapDYREL", [("Xc", "c"), ("Yc", "c"), ("X0c_ex", "E"), ("Y0c", "c")])
GershgorinPoisson = _make("GershgorinPoisson", "o_GershgorinPoisson",
                          [("G", "c"), ("D", "c"), ("rhoC", "c"), ("kx", "Fx"), ("ky", "Fy"), ("kz", "Fz"),
                           ("transient", "d"), ("dt", "d"), ("dx", "d"), ("dy", "d"), ("dz", "d")])
GershgorinHydro = _make("GershgorinHydro", "o_GershgorinHydro",
                        [("G", "c"), ("phidrhodP", "c"), ("kx", "Fx"), ("ky", "Fy"), ("kz", "Fz"), ("transient", "d"),
                         ("dt", "d"), ("dx", "d"), ("dy", "d"), ("dz", "d")])
RayleighQuotientNumerator = _make("RayleighQuotientNumerator", "o_RayleighQuotientNumerator",
                                  [("G", "c"), ("Fc0", "c"), ("Fcit", "c"), ("Fc", "c"), ("dtau", "d")])
RayleighQuotientDenominator = _make("RayleighQuotientDenominator", "o_RayleighQuotientDenominator",
                                    [("G", "c"), ("Fc0", "c"), ("dtau", "d")])
ResidualTemperatureLinearised = _make(
    "ResidualTemperatureLinearised", "o_ResidualTemperatureLinearised",
    [("F", "c"), ("Tc_ex", "E"), ("Tc_old", "c"), ("rhoC_eff", "c"), ("phc_ex", "E"), ("PC", "c"), ("qx", "Fx"),
     ("qy", "Fy"), ("qz", "Fz"), ("Qt", "d"), ("transient", "d"), ("_dt", "d"), ("_dx", "d"), ("_dy", "d"), ("_dz", "d")])
ResidualFluidPressure = _make(
    "ResidualFluidPressure", "o_ResidualFluidPressure",
    [("F", "c"), ("phc_ex", "E"), ("phirhof", "c"), ("phirhof0", "c"), ("PC", "c"), ("qx", "Fx"), ("qy", "Fy"),
     ("qz", "Fz"), ("transient", "d"), ("_dt", "d"), ("_dx", "d"), ("_dy", "d"), ("_dz", "d")])
ComputeDarcyFlux = _make(
    "ComputeDarcyFlux", "o_ComputeDarcyFlux",
    [("qx", "Fx"), ("qy", "Fy"), ("qz", "Fz"), ("rhofv", "V"), ("kx", "Fx"), ("ky", "Fy"), ("kz", "Fz"), ("Pf", "E"),
     ("g", "d"), ("_dx", "d"), ("_dy", "d"), ("_dz", "d")])
Init_vel = _make("Init_vel", "o_Init_vel",
                 [("Vx", "Fx"), ("Vy", "Fy"), ("Vz", "Fz"), ("kx", "Fx"), ("ky", "Fy"), ("kz", "Fz"), ("Pc_ex", "E"),
                  ("Ty", "Fy"), ("Ra", "d"), ("_dx", "d"), ("_dy", "d"), ("_dz", "d")])
VxPlusMinus = _make("VxPlusMinus", "o_VxPlusMinus", [("Vxm", "c"), ("Vxp", "c"), ("Vx", "Fx")])
VyPlusMinus = _make("VyPlusMinus", "o_VyPlusMinus", [("Vym", "c"), ("Vyp", "c"), ("Vy", "Fy")])
VzPlusMinus = _make("VzPlusMinus", "o_VzPlusMinus", [("Vzm", "c"), ("Vzp", "c"), ("Vz", "Fz")])
dFdx_Weno5 = _make("dFdx_Weno5", "o_dFdx_Weno5",
                   [("dFdxi", "c"), ("V1", "c"), ("V2", "c"), ("V3", "c"), ("V4", "c"), ("V5", "c")])
Gradients_Weno5 = _make("Gradients_Weno5", "o_Gradients_Weno5",
                        [("v1", "c"), ("v2", "c"), ("v3", "c"), ("v4", "c"), ("v5", "c"), ("Fc_exxx", "X6"),
                         ("axis", "i"), ("plus", "i"), ("_dx", "d"), ("_dy", "d"), ("_dz", "d")])
Advect = _make("Advect", "o_Advect", [("Fc", "c"), ("Vxp", "c"), ("dTdxm", "c"), ("Vxm", "c"), ("dTdxp", "c"), ("dt", "d")])
TimeAveraging = _make("TimeAveraging", "o_TimeAveraging", [("Fc", "c"), ("Fold", "c"), ("order", "d")])
TimeAveraging_v2 = _make("TimeAveraging_v2", "o_TimeAveraging_v2", [("Fc", "E"), ("Fold", "c"), ("order", "d")])


# Dimensional (v18) property kernels and immersed-surface ghost fills (ht3d_oracle_dim.c)
ComputeThermalConductivity = _make(
    "ComputeThermalConductivity", "o_ComputeThermalConductivity",
    [("ktv", "V"), ("Tce", "E"), ("phv", "V"), ("phi0", "d"), ("scale_kt", "d"), ("scale_T", "d")])
ComputeHydroConductivity = _make(
    "ComputeHydroConductivity", "o_ComputeHydroConductivity",
    [("kfv", "V"), ("Tce", "E"), ("Pce", "E"), ("phv", "V"), ("ymin", "d"), ("dy", "d"), ("k_fact", "d"),
     ("delta", "d"), ("scale_sigma", "d"), ("scale_T", "d"), ("scale_L", "d"), ("scale_t", "d"),
     ("scale_eta", "d"), ("mode", "i")])
ComputeFluidDensity = _make(
    "ComputeFluidDensity", "o_ComputeFluidDensity",
    [("rhofv", "V"), ("Tce", "E"), ("Pce", "E"), ("phv", "V"), ("scale_sigma", "d"), ("scale_T", "d"),
     ("scale_rho", "d")])
SmoothConductivityV2C = _make("SmoothConductivityV2C", "o_SmoothConductivityV2C", [("ktc", "c"), ("ktv", "V")])
SmoothConductivityC2V = _make("SmoothConductivityC2V", "o_SmoothConductivityC2V", [("ktv", "V"), ("ktc", "c")])
ComputeRhoCeffective = _make(
    "ComputeRhoCeffective", "o_ComputeRhoCeffective",
    [("rhoC_eff", "c"), ("Tc_ex", "E"), ("Pc_ex", "E"), ("phc_ex", "E"), ("phi_i", "d"), ("sca_T", "d"),
     ("sca_sigma", "d"), ("sca_C", "d"), ("sca_rho", "d")])
_PHI = [("phc_ex", "E"), ("phi_i", "d"), ("scale_rho", "d"), ("scale_T", "d"), ("scale_sigma", "d")]
PhiRhof = _make("PhiRhof", "o_PhiRhof", [("out", "c"), ("Tc_ex", "E"), ("Pc_ex", "E")] + _PHI)
PhiRhof_old = _make("PhiRhof_old", "o_PhiRhof_old", [("out", "c"), ("Tc0", "c"), ("Pc0", "c")] + _PHI)
PhiDrhoDP = _make("PhiDrhoDP", "o_PhiDrhoDP", [("out", "c"), ("Tc_ex", "E"), ("Pc_ex", "E")] + _PHI)
dim_Init_vel = _make("dim_Init_vel", "o_dim_Init_vel",
                     [("Vx", "Fx"), ("Vy", "Fy"), ("Vz", "Fz"), ("qx", "Fx"), ("qy", "Fy"), ("qz", "Fz"),
                      ("rhofv", "V")])
dim_SetTemperatureBCs = _make(
    "dim_SetTemperatureBCs", "o_dim_SetTemperatureBCs",
    [("Tc_ex", "E"), ("ky", "Fy"), ("wN", "W"), ("phc_ex", "E")] +
    [(n, "d") for n in ("qyS", "_dy", "kS", "TS", "TN", "y_plateau", "a2", "b2", "dTdy", "dx", "dy", "sticky_air",
                        "sca_T", "sca_kT", "phi_i")])
dim_SetPressureBCs = _make(
    "dim_SetPressureBCs", "o_dim_SetPressureBCs",
    [("Pc_ex", "E"), ("wN", "W"), ("phc_ex", "E")] +
    [(n, "d") for n in ("PS", "PN", "y_plateau", "a2", "b2", "rhof", "g", "dx", "dy", "sticky_air")])


Phase = _make("Phase", "o_Phase", [("X", "c"), ("phc_ex", "E")])
Pressure = _make("Pressure", "o_Pressure", [("X", "c"), ("Pc_ex", "E"), ("phc_ex", "E"), ("scale", "d")])
Temperature = Pressure
Velocity = _make("Velocity", "o_Velocity", [("X", "c"), ("Y", "c"), ("Z", "c"), ("Vx", "Fx"), ("Vy", "Fy"),
                                            ("Vz", "Fz"), ("phc_ex", "E"), ("scale_V", "d")])
EffectiveThermalConductivity = _make("EffectiveThermalConductivity", "o_EffectiveThermalConductivity",
                                     [("X", "c"), ("Tc_ex", "E"), ("phi0", "d"), ("phc_ex", "E"), ("scale_T", "d")])
EffectiveHeatCapacity = _make("EffectiveHeatCapacity", "o_EffectiveHeatCapacity",
                              [("Ceff", "c"), ("Tc_ex", "E"), ("Pc_ex", "E"), ("phc_ex", "E")] +
                              [(n, "d") for n in ("phi_i", "sca_T", "sca_sigma", "sca_C", "sca_rho")])
Permeability = _make("Permeability", "o_Permeability", [("X", "c"), ("k_fact", "d"), ("ymin", "d"), ("dy", "d"),
                                                        ("phc_ex", "E"), ("delta", "d"), ("scale_L", "d")])
FluidDensity = _make("FluidDensity", "o_FluidDensity", [("X", "c"), ("Tc_ex", "E"), ("Pc_ex", "E"), ("phc_ex", "E"),
                                                        ("scale_T", "d"), ("scale_sigma", "d")])
Viscosity = _make("Viscosity", "o_Viscosity", [("X", "c"), ("Tc_ex", "E"), ("phc_ex", "E"), ("scale_T", "d")])
Peclet = _make("Peclet", "o_Peclet", [("X", "c"), ("Tc_ex", "E"), ("Pc_ex", "E"), ("Vx", "Fx"), ("Vy", "Fy"),
                                      ("Vz", "Fz"), ("phc_ex", "E")] +
               [(n, "d") for n in ("phi0", "scale_T", "scale_sigma", "scale_V", "scale_L")])

EOS = {"kTs": 0, "kTf": 1, "Cf": 2, "Cs": 3, "rhos": 4, "rhof_C": 5, "drhodP_C": 6, "muf": 7, "kF": 8, "phi": 9,
       "pow4": 10}


def eos(name: str, a: float, b: float = 0.0) -> float:
    """The oracle's scalar EOS / property closure `name` (HydroThermal3D_v18.jl:29-39)."""
    return lib().o_eos(EOS[name], C.c_double(a), C.c_double(b))


def Boundaries_Weno5(Fc_exxx, Fc, axis, type_lo, val_lo, type_hi, val_hi):
    grid = _grid(Fc_exxx, "X6")
    src_ext = 1 if Fc.shape == tuple(g + 2 for g in grid) else 0
    lib().o_Boundaries_Weno5(_ptr(Fc_exxx), _ptr(Fc), int(axis), src_ext, int(type_lo), C.c_double(val_lo),
                             int(type_hi), C.c_double(val_hi), *grid)


def Advect_v2(Fc, X, V, axis, dt, _dx, _dy, _dz):
    grid = _grid(Fc, "E")
    lib().o_Advect_v2(_ptr(Fc), _ptr(X), _ptr(V), int(axis), C.c_double(dt), C.c_double(_dx), C.c_double(_dy),
                      C.c_double(_dz), *grid)


def AdvectWithWeno5_v2(Tc_ex, Tc_exxx, Told, Vx, Vy, Vz, dt, _dx, _dy, _dz, Ttop, Tbot):
    """tools/Weno5_Routines_v2.jl:1-50 (C orchestration of the v2 kernels)."""
    grid = _grid(Tc_ex, "E")
    lib().o_AdvectWithWeno5_v2(_ptr(Tc_ex), _ptr(Tc_exxx), _ptr(Told), _ptr(Vx), _ptr(Vy), _ptr(Vz), C.c_double(dt),
                               C.c_double(_dx), C.c_double(_dy), C.c_double(_dz), C.c_double(Ttop),
                               C.c_double(Tbot), *grid)


def AdvectWithWeno5(Tc_ex, Vx, Vy, Vz, dt, _dx, _dy, _dz, Ttop, Tbot):
    """Dimensionless/kernels_HT3D.jl:156-233 after Init_vel!, orchestrated here from
    the oracle's v1 kernels in the reference's launch order."""
    nx, ny, nz = _grid(Tc_ex, "E")
    c = (nx, ny, nz)
    z = lambda: np.zeros(c, order="F")
    Tc, Told, v = z(), z(), [z() for _ in range(5)]
    dTdxm, dTdxp = z(), z()
    Vm = {a: z() for a in "xyz"}
    Vp = {a: z() for a in "xyz"}
    X6 = np.zeros((nx + 6, ny + 6, nz + 6), order="F")
    ResetA(Vm["x"], Vp["x"]); VxPlusMinus(Vm["x"], Vp["x"], Vx)
    ResetA(Vm["y"], Vp["y"]); VyPlusMinus(Vm["y"], Vp["y"], Vy)
    ResetA(Vm["z"], Vp["z"]); VzPlusMinus(Vm["z"], Vp["z"], Vz)
    Cpy_inn_to_all(Tc, Tc_ex)
    bcs = {1: (0, 0.0, 0, 0.0), 2: (1, Tbot, 1, Ttop), 3: (0, 0.0, 0, 0.0)}
    for axis, a in ((1, "x"), (2, "y"), (3, "z")):
        ArrayEqualArray(Told, Tc)
        for _ in range(2):
            Boundaries_Weno5(X6, Tc, axis, *bcs[axis])
            Gradients_Weno5(*v, X6, axis, 0, _dx, _dy, _dz)
            dFdx_Weno5(dTdxm, *v)
            Gradients_Weno5(*v, X6, axis, 1, _dx, _dy, _dz)
            dFdx_Weno5(dTdxp, *v)
            Advect(Tc, Vp[a], dTdxm, Vm[a], dTdxp, dt)
        TimeAveraging(Tc, Told, 2.0)
    Cpy_all_to_inn(Tc_ex, Tc)


def weno5_js(v1, v2, v3, v4, v5) -> float:
    return lib().o_weno5_js(v1, v2, v3, v4, v5)


SUM, SUM_ABS, MAX, MIN, MAX_ABS = 0, 1, 2, 3, 4


def reduce(a: np.ndarray, op: int) -> float:
    a = np.asfortranarray(a, dtype=np.float64)
    return lib().o_reduce(_ptr(a), C.c_size_t(a.shape[0]), C.c_size_t(a.size // a.shape[0]), op)


def set_threads(n: int) -> int:
    """Set the oracle's OpenMP thread count; returns the previous one."""
    old = lib().o_get_threads()
    lib().o_set_threads(int(n))
    return old


def lognormal_field(kfv, xmin, ymin, zmin, dx, dy, dz, K, phi, M, sigma, fault_mult, fault_half, n, x0):
    nx, ny, nz = (s - 1 for s in kfv.shape)
    K = np.ascontiguousarray(K, dtype=np.float64)
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    n = np.ascontiguousarray(n, dtype=np.float64)
    x0 = np.ascontiguousarray(x0, dtype=np.float64)
    lib().o_lognormal_field(_ptr(kfv), nx, ny, nz, C.c_double(xmin), C.c_double(ymin), C.c_double(zmin),
                            C.c_double(dx), C.c_double(dy), C.c_double(dz), C.c_void_p(K.ctypes.data),
                            C.c_void_p(phi.ctypes.data), int(M), C.c_double(sigma), C.c_double(fault_mult),
                            C.c_double(fault_half), C.c_void_p(n.ctypes.data), C.c_void_p(x0.ctypes.data))


# --------------------------------------------------------------- DYREL / model
class Params(C.Structure):
    _fields_ = [("nitmax", C.c_int), ("nitout", C.c_int), ("ncheck", C.c_int), ("norm", C.c_int),
                ("tol", C.c_double), ("CFL", C.c_double), ("cfact", C.c_double), ("lmin_div", C.c_double)]


class Report(C.Structure):
    _fields_ = [("iters", C.c_int), ("status", C.c_int), ("n_checks", C.c_int), ("n_lambda", C.c_int),
                ("lmax", C.c_double), ("lmin", C.c_double), ("dtau", C.c_double), ("h1", C.c_double),
                ("h2", C.c_double), ("nF_abs", C.c_double), ("nF_rel", C.c_double), ("nF_ini", C.c_double)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class Field(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("X", C.c_void_p), ("R", C.c_void_p),
                ("kx", C.c_void_p), ("ky", C.c_void_p), ("kz", C.c_void_p), ("PC", C.c_void_p), ("G", C.c_void_p),
                ("F", C.c_void_p), ("F0", C.c_void_p), ("Fcit", C.c_void_p), ("dumc", C.c_void_p),
                ("qx", C.c_void_p), ("qy", C.c_void_p), ("qz", C.c_void_p), ("transient", C.c_int),
                ("rhoC", C.c_double), ("dt", C.c_double), ("_dt", C.c_double), ("dx", C.c_double),
                ("dy", C.c_double), ("dz", C.c_double), ("bc_type", C.c_int * 6), ("bc_val", C.c_double * 6)]


def params(nitmax=10000, nitout=1000, ncheck=0, norm=0, tol=1e-8, CFL=0.99, cfact=0.5, lmin_div=500.0) -> Params:
    return Params(nitmax, nitout, ncheck, norm, tol, CFL, cfact, lmin_div)


def dyrel_solve(X, R, kx, ky, kz, *, dx, dy, dz, dt, transient, rhoC, bc_type=(0,) * 6, bc_val=(0.0,) * 6,
                p: Params | None = None) -> tuple[dict, dict]:
    """Run the unfused DYREL solve; returns (report, work arrays)."""
    nx, ny, nz = _grid(X, "E")
    shp = (nx, ny, nz)
    work = {k: np.zeros(shp, order="F") for k in ("PC", "G", "F", "F0", "Fcit", "dumc")}
    work["qx"] = np.zeros((nx + 1, ny, nz), order="F")
    work["qy"] = np.zeros((nx, ny + 1, nz), order="F")
    work["qz"] = np.zeros((nx, ny, nz + 1), order="F")
    f = Field(nx, ny, nz, _ptr(X), _ptr(R), _ptr(kx), _ptr(ky), _ptr(kz), *(
        _ptr(work[k]) for k in ("PC", "G", "F", "F0", "Fcit", "dumc", "qx", "qy", "qz")), int(transient),
        float(rhoC), float(dt), 1.0 / dt, dx, dy, dz, (C.c_int * 6)(*bc_type), (C.c_double * 6)(*bc_val))
    rep = Report()
    lib().o_dyrel_solve(C.byref(f), C.byref(p or params()), C.byref(rep))
    return rep.as_dict(), work


class HRLModel:
    """The oracle's closed-box HRL model (o_hrl_*), stepping on the CPU."""

    SHAPES = {"Tc_ex": "E", "Pc_ex": "E", "ktv": "V", "kfv": "V", "Ty": "Fy", "kx": "Fx", "ky": "Fy", "kz": "Fz",
              "qx": "Fx", "qy": "Fy", "qz": "Fz", "Vx": "Fx", "Vy": "Fy", "Vz": "Fz", "Rt": "c", "Rp": "c", "F": "c",
              "F0": "c", "Fcit": "c", "PC": "c", "G": "c"}

    def __init__(self, nx, ny, nz, Lx, Ly, Lz, Ra, cfl=0.5):
        self._m = lib().o_hrl_create(nx, ny, nz, Lx, Ly, Lz, Ra, cfl)
        self.grid = (nx, ny, nz)

    def field(self, name: str) -> np.ndarray:
        """A live view of the model's array (Fortran order)."""
        p = lib().o_hrl_field(self._m, name.encode())
        shape = tuple(g + d for g, d in zip(self.grid, _OFF[self.SHAPES[name]]))
        buf = (C.c_double * int(np.prod(shape))).from_address(p)
        return np.ndarray(shape, dtype=np.float64, buffer=buf, order="F")

    def set_params(self, pT: Params, pP: Params, weno_version: int = 1):
        lib().o_hrl_set_params(self._m, C.byref(pT), C.byref(pP), weno_version)

    def set_dt(self, dt: float):
        lib().o_hrl_set_dt(self._m, C.c_double(dt))

    def set_viscosity(self, a: float):
        lib().o_hrl_set_viscosity(self._m, C.c_double(a))

    def set_dt_rule(self, rule: int):
        """0: v18 (advect with dt_adv), 1: v18 constant dt with sub-steps, 2: v7."""
        lib().o_hrl_set_dt_rule(self._m, int(rule))

    def step(self) -> dict:
        st = lib().o_hrl_step(self._m)
        rT, rP = Report(), Report()
        lib().o_hrl_get_reports(self._m, C.byref(rT), C.byref(rP))
        return {"status": st, "thermal": rT.as_dict(), "darcy": rP.as_dict(), "dt_next": self.scalar(0),
                "time": self.scalar(1), "Nu": self.scalar(2), "n_adv": lib().o_hrl_n_adv(self._m)}

    def scalar(self, which: int) -> float:
        return lib().o_hrl_scalar(self._m, which)

    def close(self):
        if self._m:
            lib().o_hrl_destroy(self._m)
            self._m = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# --------------------------------------------------------------- v18 DYREL solves (ht3d_oracle_dim.c)
class V18Thermal(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("dx", C.c_double), ("dy", C.c_double),
                ("dz", C.c_double), ("dt", C.c_double), ("transient", C.c_double), ("Tc_ex", C.c_void_p),
                ("Tc0", C.c_void_p), ("rhoC", C.c_void_p), ("kx", C.c_void_p), ("ky", C.c_void_p),
                ("kz", C.c_void_p), ("phc_ex", C.c_void_p), ("wN", C.c_void_p), ("qyS", C.c_double),
                ("TN", C.c_double), ("Qt", C.c_double), ("norm_scale", C.c_double)] + \
               [(n, C.c_void_p) for n in ("G", "PC", "F", "F0", "Fcit", "dumc", "qx", "qy", "qz")]


class V18Darcy(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("dx", C.c_double), ("dy", C.c_double),
                ("dz", C.c_double), ("dt", C.c_double), ("transient", C.c_double), ("Pc_ex", C.c_void_p),
                ("Tc_ex", C.c_void_p), ("phirho", C.c_void_p), ("phirhof0", C.c_void_p), ("rhofv", C.c_void_p),
                ("kx", C.c_void_p), ("ky", C.c_void_p), ("kz", C.c_void_p), ("phc_ex", C.c_void_p),
                ("wN", C.c_void_p), ("PN", C.c_double), ("g", C.c_double), ("phi_i", C.c_double),
                ("scale_rho", C.c_double), ("scale_T", C.c_double), ("scale_sigma", C.c_double),
                ("norm_scale", C.c_double)] + \
               [(n, C.c_void_p) for n in ("G", "PC", "F", "F0", "Fcit", "dumc", "qx", "qy", "qz")]


def _v18_work(nx, ny, nz):
    w = {k: np.zeros((nx, ny, nz), order="F") for k in ("G", "PC", "F", "F0", "Fcit", "dumc")}
    w["qx"] = np.zeros((nx + 1, ny, nz), order="F")
    w["qy"] = np.zeros((nx, ny + 1, nz), order="F")
    w["qz"] = np.zeros((nx, ny, nz + 1), order="F")
    return w


def v18_thermal_solve(Tc_ex, Tc0, rhoC, kx, ky, kz, phc_ex, wN, *, dx, dy, dz, dt, transient, qyS, TN, Qt=0.0,
                      norm_scale=1.0, p: Params | None = None):
    """HydroThermal3D_v18.jl:432-478, unfused; returns (report, work arrays)."""
    nx, ny, nz = _grid(Tc_ex, "E")
    w = _v18_work(nx, ny, nz)
    f = V18Thermal(nx, ny, nz, dx, dy, dz, dt, float(transient), _ptr(Tc_ex), _ptr(Tc0), _ptr(rhoC), _ptr(kx),
                   _ptr(ky), _ptr(kz), _ptr(phc_ex), _ptr(wN), qyS, TN, Qt, norm_scale,
                   *(_ptr(w[k]) for k in ("G", "PC", "F", "F0", "Fcit", "dumc", "qx", "qy", "qz")))
    rep = Report()
    lib().o_v18_thermal_solve(C.byref(f), C.byref(p or params()), C.byref(rep))
    return rep.as_dict(), w


def v18_darcy_solve(Pc_ex, Tc_ex, phirho, phirhof0, rhofv, kx, ky, kz, phc_ex, wN, *, dx, dy, dz, dt, transient, PN,
                    g, phi_i, scale_rho, scale_T, scale_sigma, norm_scale=1.0, p: Params | None = None):
    """HydroThermal3D_v18.jl:489-537, unfused; returns (report, work arrays)."""
    nx, ny, nz = _grid(Pc_ex, "E")
    w = _v18_work(nx, ny, nz)
    f = V18Darcy(nx, ny, nz, dx, dy, dz, dt, float(transient), _ptr(Pc_ex), _ptr(Tc_ex), _ptr(phirho),
                 _ptr(phirhof0), _ptr(rhofv), _ptr(kx), _ptr(ky), _ptr(kz), _ptr(phc_ex), _ptr(wN), PN, g, phi_i,
                 scale_rho, scale_T, scale_sigma, norm_scale,
                 *(_ptr(w[k]) for k in ("G", "PC", "F", "F0", "Fcit", "dumc", "qx", "qy", "qz")))
    rep = Report()
    lib().o_v18_darcy_solve(C.byref(f), C.byref(p or params()), C.byref(rep))
    return rep.as_dict(), w
