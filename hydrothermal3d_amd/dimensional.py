"""Characteristic scales of the dimensional (v18) model.

Also the fused DYREL solves of the v18 model (`v18_thermal_solve`,
`v18_darcy_solve`): the heat and Darcy loops of
Dimensional/HydroThermal3D_v18.jl:432-478 and :489-537.

`Scaling` and `scale_me` mirror the `scaling` struct and `scale_me!` of
Dimensional/kernels_HT3D_dimensional.jl:3-36: four primary scales (T, V, L, eta)
set by the driver, and the derived ones (time, stress, mass, density, heat
capacity, thermal conductivity, permeability ...) that the property kernels
take as arguments.  `v18_scaling()` returns the values of
Dimensional/HydroThermal3D_v18.jl:137-142.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

from . import _lib
from .device import DeviceArray


@dataclass
class Scaling:
    T: float = float("nan")
    V: float = float("nan")
    L: float = float("nan")
    eta: float = float("nan")
    t: float = float("nan")
    a: float = float("nan")
    eps: float = float("nan")
    sigma: float = float("nan")
    m: float = float("nan")
    rho: float = float("nan")
    F: float = float("nan")
    J: float = float("nan")
    W: float = float("nan")
    C: float = float("nan")
    kt: float = float("nan")
    kf: float = float("nan")


def scale_me(s: Scaling) -> Scaling:
    """Derived scales, in the order and with the formulas of scale_me!
    (x^2 = x*x and x^3 = x*x*x, as Julia evaluates literal powers)."""
    s.t = s.L / s.V
    s.a = s.V / s.t
    s.eps = 1.0 / s.t
    s.sigma = s.eta / s.t
    s.m = s.sigma * s.L * (s.t * s.t)
    s.rho = s.m / (s.L * s.L * s.L)
    s.F = s.m * s.L / (s.t * s.t)
    s.J = s.m * (s.L * s.L) / (s.t * s.t)
    s.W = s.J / s.t
    s.C = s.J / s.m / s.T
    s.kt = s.W / s.L / s.T
    s.kf = s.L * s.L
    return s


def v18_scaling() -> Scaling:
    return scale_me(Scaling(T=500.0, V=1e-9, L=700.0, eta=1e10))


def _v18_call(name: str, X: DeviceArray, shapes: dict, arrays: dict, scalars: dict, params) -> dict:
    from .kernels import _check, _grid_from
    from .solver import DyrelParams, _status_check
    grid = _grid_from(X, "E")
    for nm, (a, kind) in shapes.items():
        if a is None:
            continue
        if kind == "W":
            if not isinstance(a, DeviceArray) or a.shape != (grid[0] + 2, grid[2] + 2):
                raise ValueError(f"{name}: {nm} must be a DeviceArray of shape {(grid[0] + 2, grid[2] + 2)}")
        else:
            _check(a, kind, grid, name, nm)
    f = _lib.V18Field(grid[0], grid[1], grid[2])
    f.X = X.ptr
    for nm, a in arrays.items():
        setattr(f, nm, a.ptr if a is not None else None)
    for nm, v in scalars.items():
        setattr(f, nm, float(v))
    p = (params or DyrelParams()).c()
    rep = _lib.DyrelReport()
    _lib.ensure_init()
    st = getattr(_lib.lib(), "ht3d_" + name)(C.byref(f), C.byref(p), C.byref(rep))
    if st not in (_lib.OK, _lib.NOT_CONVERGED, _lib.NAN, _lib.DIVERGED):
        _lib.check(st, "ht3d_" + name)
    _status_check(st, "T" if "thermal" in name else "P", rep.as_dict())
    return rep.as_dict()


def v18_thermal_solve(Tc_ex, Tc0, rhoC, kx, ky, kz, phc_ex, wN, *, dx, dy, dz, dt, transient, qyS, TN, Qt=0.0,
                      norm_scale=1.0, params=None, PC=None, G=None, kv=None) -> dict:
    """The heat DYREL loop of HydroThermal3D_v18.jl:432-478 on the device (one
    fused launch per pseudo-iteration): Tc_ex is updated in place, PC/G (if
    given) receive the Gershgorin diagonal and bound.  kv: the vertex field
    whose face averages kx, ky, kz are (InitConductivity!), which the tuner may
    then read instead of the three face arrays (same results bit for bit).
    Returns the report."""
    shapes = {"Tc0": (Tc0, "c"), "rhoC": (rhoC, "c"), "kx": (kx, "Fx"), "ky": (ky, "Fy"), "kz": (kz, "Fz"),
              "phc_ex": (phc_ex, "E"), "wN": (wN, "W"), "PC": (PC, "c"), "G": (G, "c"), "kv": (kv, "V")}
    arrays = {"Tc_old": Tc0, "rhoC": rhoC, "kx": kx, "ky": ky, "kz": kz, "phc_ex": phc_ex, "wN": wN, "PC": PC,
              "G": G, "kv": kv}
    scalars = {"dx": dx, "dy": dy, "dz": dz, "dt": dt, "transient": transient, "qyS": qyS, "top": TN, "Qt": Qt,
               "norm_scale": norm_scale}
    return _v18_call("v18_thermal_solve", Tc_ex, shapes, arrays, scalars, params)


def v18_darcy_solve(Pc_ex, Tc_ex, phirho, phirhof0, rhofv, kx, ky, kz, phc_ex, wN, *, dx, dy, dz, dt, transient,
                    PN, g, phi_i, scale_rho, scale_T, scale_sigma, norm_scale=1.0, params=None, PC=None,
                    G=None, kv=None) -> dict:
    """The Darcy DYREL loop of HydroThermal3D_v18.jl:489-537 on the device:
    phirho holds phi*drho/dP on entry (the Gershgorin diagonal) and phi*rho_f of
    the last iteration on exit, as the reference's shared buffer does.  kv: as
    for v18_thermal_solve (the hydraulic conductivity on vertices)."""
    shapes = {"Tc_ex": (Tc_ex, "E"), "phirho": (phirho, "c"), "phirhof0": (phirhof0, "c"), "rhofv": (rhofv, "V"),
              "kx": (kx, "Fx"), "ky": (ky, "Fy"), "kz": (kz, "Fz"), "phc_ex": (phc_ex, "E"), "wN": (wN, "W"),
              "PC": (PC, "c"), "G": (G, "c"), "kv": (kv, "V")}
    arrays = {"Tc_ex": Tc_ex, "rhoC": phirho, "phirhof0": phirhof0, "rhofv": rhofv, "kx": kx, "ky": ky, "kz": kz,
              "phc_ex": phc_ex, "wN": wN, "PC": PC, "G": G, "kv": kv}
    scalars = {"dx": dx, "dy": dy, "dz": dz, "dt": dt, "transient": transient, "top": PN, "g": g, "phi_i": phi_i,
               "scale_rho": scale_rho, "scale_T": scale_T, "scale_sigma": scale_sigma, "norm_scale": norm_scale}
    return _v18_call("v18_darcy_solve", Pc_ex, shapes, arrays, scalars, params)
