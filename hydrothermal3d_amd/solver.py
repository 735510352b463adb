"""Fused solves and the closed-box physical step on the device.

`dyrel_solve` is one DYREL solve of HydroThermal3D_v18.jl:424-478 (:481-538) on
the dimensionless operators; `HydroThermal3D` steps closed-box Horton-Rogers-
Lapwood convection (Dimensionless/HydroThermal3D_Alban_v6.jl) with DYREL solves
and WENO5 advection, every kernel on the GPU.  Solver failures that the
reference turns into `error(...)` (v18:464-465) raise here; a solve that hits
nitmax is reported (the reference is silent about it, v7:173-184).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, asdict

import numpy as np

from . import _lib
from .device import DeviceArray
from .kernels import _check, _grid_from


@dataclass
class DyrelParams:
    """DYREL controls; defaults are those of HydroThermal3D_v18.jl:309-310,434-436."""
    nitmax: int = 10000
    nitout: int = 1000       # lambda_min refresh + reference check: iter<=2 or iter % nitout == 0
    ncheck: int = 0          # > 0: also test convergence every ncheck iterations
    norm: int = 0            # 0: mean|F*D|/sqrt(N) (v18:459-460); 1: max|F*D|
    tol: float = 1e-8
    CFL: float = 0.99
    cfact: float = 0.5
    lmin_div: float = 500.0

    def c(self) -> _lib.DyrelParams:
        return _lib.DyrelParams(self.nitmax, self.nitout, self.ncheck, self.norm, self.tol, self.CFL, self.cfact,
                                self.lmin_div)


def reference_params(tol: float = 1e-8) -> DyrelParams:
    """The reference schedule: lambda_min and convergence only at iter<=2 or every 1000."""
    return DyrelParams(tol=tol)


def production_params(tol: float = 1e-8, nitout: int = 100, ncheck: int = 50, nitmax: int = 100000) -> DyrelParams:
    """Max-norm convergence test every `ncheck` iterations, lambda_min refreshed
    every `nitout` (the configs' stopping rule: tolerance on max|F*D|)."""
    return DyrelParams(nitmax=nitmax, nitout=nitout, ncheck=ncheck, norm=1, tol=tol)


class SolveError(RuntimeError):
    """A solve stopped on NaN or divergence (the reference's error() calls);
    `report` holds the solve report when the caller has one."""

    def __init__(self, msg: str, report: dict | None = None):
        super().__init__(msg)
        self.report = report


def _step_dict(rep: _lib.StepReport) -> dict:
    out = {"thermal": rep.thermal.as_dict(), "darcy": rep.darcy.as_dict()}
    for name, _ in _lib.StepReport._fields_[2:]:
        out[name] = getattr(rep, name)
    return out


def _status_check(status: int, what: str, report: dict | None = None) -> None:
    if status == _lib.NAN:
        raise SolveError(f"Nan {what}...", report)          # v18:464
    if status == _lib.DIVERGED:
        raise SolveError(f"Diverged {what}...", report)     # v18:465


def dyrel_solve(X: DeviceArray, R: DeviceArray, kx: DeviceArray, ky: DeviceArray, kz: DeviceArray, *,
                dx: float, dy: float, dz: float, dt: float, transient: int, rhoC: float,
                bc_type=(0, 0, 0, 0, 0, 0), bc_val=(0.0,) * 6, params: DyrelParams | None = None,
                precision: int = _lib.FP64, kv: DeviceArray | None = None) -> dict:
    """One DYREL solve on the device; returns the solve report as a dict.

    kv: the vertex field whose face averages kx, ky, kz are (InitThermal! /
    InitDarcy!, kernels_HT3D.jl:82-107); the FP64 solve may then read it
    instead of the three face arrays, with the same results bit for bit."""
    grid = _grid_from(X, "E")
    for nm, a, k in (("X", X, "E"), ("R", R, "c"), ("kx", kx, "Fx"), ("ky", ky, "Fy"), ("kz", kz, "Fz")):
        _check(a, k, grid, "dyrel_solve", nm)
    if kv is not None:
        _check(kv, "V", grid, "dyrel_solve", "kv")
    p = (params or DyrelParams()).c()
    f = _lib.DyrelField(grid[0], grid[1], grid[2], dx, dy, dz, X.ptr, R.ptr, kx.ptr, ky.ptr, kz.ptr, int(transient),
                        float(rhoC), float(dt), (C.c_int * 6)(*bc_type), (C.c_double * 6)(*bc_val),
                        kv.ptr if kv is not None else None)
    rep = _lib.DyrelReport()
    _lib.ensure_init()
    if precision == _lib.FP64:
        st = _lib.lib().ht3d_dyrel_solve(C.byref(f), C.byref(p), C.byref(rep))
        out = rep.as_dict()
    else:
        outer, inner = C.c_int(), C.c_int()
        st = _lib.lib().ht3d_dyrel_solve_dc(C.byref(f), C.byref(p), C.byref(rep), C.byref(outer), C.byref(inner))
        out = rep.as_dict()
        out.update(dc_outer=outer.value, dc_inner=inner.value)
    if st not in (_lib.OK, _lib.NOT_CONVERGED, _lib.NAN, _lib.DIVERGED):
        _lib.check(st, "ht3d_dyrel_solve")
    _status_check(st, "T" if transient else "P", out)
    return out


class HydroThermal3D:
    """Closed-box HRL porous convection on one GPU (reference axes: y vertical).

    Fields are device arrays shaped like the reference's (Tc_ex, Pc_ex: E;
    ktv, kfv: V; Vx/Vy/Vz: faces).  `step()` runs thermal solve, Nu, Darcy solve,
    velocity, the advective CFL time step from that velocity and WENO5
    advection (HydroThermal3D_v18.jl:551-570), all on the device.
    """

    FIELDS = {"Tc_ex": "E", "Pc_ex": "E", "ktv": "V", "kfv": "V", "Ty": "Fy", "kx": "Fx", "ky": "Fy", "kz": "Fz",
              "ktx": "Fx", "kty": "Fy", "ktz": "Fz", "Vx": "Fx", "Vy": "Fy", "Vz": "Fz", "Rt": "c", "Rp": "c"}

    def __init__(self, ncx: int, ncy: int, ncz: int, Lx: float, Ly: float, Lz: float, Ra: float, cfl: float = 0.5):
        _lib.ensure_init()
        h = C.c_void_p()
        _lib.call("ht3d_hrl_create", ncx, ncy, ncz, C.c_double(Lx), C.c_double(Ly), C.c_double(Lz), C.c_double(Ra),
                  C.c_double(cfl), C.byref(h))
        self._h = h
        self.grid = (ncx, ncy, ncz)
        self.L = (Lx, Ly, Lz)
        self.Ra = Ra
        self._fields: dict[str, DeviceArray] = {}
        off = {"c": (0, 0, 0), "E": (2, 2, 2), "V": (1, 1, 1), "Fx": (1, 0, 0), "Fy": (0, 1, 0), "Fz": (0, 0, 1)}
        for name, kind in self.FIELDS.items():
            ptr, cnt = C.c_void_p(), C.c_size_t()
            _lib.call("ht3d_hrl_field", self._h, name.encode(), C.byref(ptr), C.byref(cnt))
            shape = tuple(g + d for g, d in zip(self.grid, off[kind]))
            self._fields[name] = DeviceArray(shape, ptr=ptr.value, owner=self)
        self.history: list[tuple[float, float]] = []   # (time, Nu) after each step
        self.set_params()

    def __getitem__(self, name: str) -> DeviceArray:
        return self._fields[name]

    def set_params(self, thermal: DyrelParams | None = None, darcy: DyrelParams | None = None,
                   weno_version: int = 1, precision: int = _lib.FP64) -> None:
        self.params = (thermal or DyrelParams(), darcy or DyrelParams(), weno_version, precision)
        pT, pP = self.params[0].c(), self.params[1].c()
        _lib.call("ht3d_hrl_set_params", self._h, C.byref(pT), C.byref(pP), weno_version, precision)

    def set_dt(self, dt: float) -> None:
        _lib.call("ht3d_hrl_set_dt", self._h, C.c_double(dt))

    def set_dt_rule(self, rule: int) -> None:
        """Time step rule: DT_V18 (advect with dt_adv = cfl h/max|V| and keep it,
        HydroThermal3D_v18.jl:553-562), DT_V18_CONSTANT (keep dt, advect in
        ceil(dt/dt_adv) sub-steps, v18:563-570) or DT_V7 (advect with dt, then
        dt = dt_adv, HydroThermal3D_v7.jl:215-221)."""
        _lib.call("ht3d_hrl_set_dt_rule", self._h, int(rule))

    def set_viscosity(self, a: float) -> None:
        """mu = exp(-a T) in the Darcy solve (a = ln 10 for the 768^3-class workload)."""
        _lib.call("ht3d_hrl_set_viscosity", self._h, C.c_double(a))

    def step(self) -> dict:
        rep = _lib.StepReport()
        st = _lib.lib().ht3d_hrl_step(self._h, C.byref(rep))
        if st not in (_lib.OK, _lib.NOT_CONVERGED, _lib.NAN, _lib.DIVERGED):
            _lib.check(st, "ht3d_hrl_step")
        _status_check(rep.thermal.status, "T")
        _status_check(rep.darcy.status, "P")
        self.history.append((rep.time, rep.Nu))
        return _step_dict(rep)

    def save_nu_history(self, name, nt: int | None = None):
        """HT3DOutput%05d.h5 (Dimensionless/HydroThermal3D_Alban_v6.jl:500-505):
        the vectors time_vec and Nu_vec of the reference, entry 1 the initial
        state (0, 0), entry it+1 the time and Nu after step it (:466-469), zero
        past the steps taken when the run's step count `nt` is given."""
        from .h5io import write_nu_history
        n = max(len(self.history) + 1, (nt + 1) if nt is not None else 0)
        t, nu = np.zeros(n), np.zeros(n)
        for q, (a, b) in enumerate(self.history):
            t[q + 1], nu[q + 1] = a, b
        return write_nu_history(name, t, nu)

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            for a in self._fields.values():
                a.ptr = 0
            _lib.lib().ht3d_hrl_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HydroThermal3DBatch:
    """Many closed-box problems stepped together (ht3d_batch_*).

    `models` are HydroThermal3D instances of any sizes, Rayleigh numbers,
    permeabilities, viscosities and time step rules; they keep their fields and
    are stepped in lockstep.  The DYREL parameters and the precision are the
    batch's; the solves of all problems run as one persistent launch per
    pseudo-iteration over a packed work list, with the runs between checks
    replayed as HIP graphs when `graphs` is true.
    """

    def __init__(self, models, thermal: DyrelParams | None = None, darcy: DyrelParams | None = None,
                 precision: int = _lib.FP64, graphs: bool = True):
        _lib.ensure_init()
        self.models = list(models)
        arr = (C.c_void_p * len(self.models))(*[m._h.value for m in self.models])
        h = C.c_void_p()
        _lib.call("ht3d_batch_create", arr, len(self.models), C.byref(h))
        self._h = h
        self.set_params(thermal, darcy, precision, graphs)

    def set_params(self, thermal: DyrelParams | None = None, darcy: DyrelParams | None = None,
                   precision: int = _lib.FP64, graphs: bool = True) -> None:
        pT, pP = (thermal or DyrelParams()).c(), (darcy or DyrelParams()).c()
        _lib.call("ht3d_batch_set_params", self._h, C.byref(pT), C.byref(pP), precision, int(bool(graphs)))

    def step(self) -> list[dict]:
        reps = (_lib.StepReport * len(self.models))()
        st = _lib.lib().ht3d_batch_step(self._h, reps)
        if st not in (_lib.OK, _lib.NOT_CONVERGED, _lib.NAN, _lib.DIVERGED):
            _lib.check(st, "ht3d_batch_step")
        for m, r in zip(self.models, reps):
            _status_check(r.thermal.status, "T")
            _status_check(r.darcy.status, "P")
            m.history.append((r.time, r.Nu))
        return [_step_dict(r) for r in reps]

    def stats(self) -> dict:
        s = _lib.BatchStats()
        _lib.call("ht3d_batch_stats_get", self._h, C.byref(s))
        return s.as_dict()

    def reset_stats(self) -> None:
        _lib.call("ht3d_batch_stats_reset", self._h)

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            _lib.lib().ht3d_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
