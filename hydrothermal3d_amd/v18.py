"""The dimensional (v18) hydrothermal model stepped on the GPU.

`V18Model.step()` is one pass of the time loop of
Dimensional/HydroThermal3D_v18.jl:398-571 with every kernel on the device:

  1. InitThermal!(Tc0, Tc_ex), InitThermal!(Pc0, Pc_ex)               (:419-420)
  2. heat: ComputeThermalConductivity, SmoothConductivityV2C/C2V,
     InitConductivity!, ComputeρCeffective!, then the DYREL heat loop  (:424-478)
  3. Darcy: ComputeHydroConductivity (mode 1), smoothing, InitConductivity!,
     ϕdρdP!, ϕρf_old, ComputeFluidDensity, then the DYREL Darcy loop     (:481-538)
  4. velocity: ComputeHydroConductivity (mode 2), smoothing,
     InitConductivity!, ComputeFluidDensity, ComputeDarcyFlux!           (:542-548)
  5. for it > 0: advection with AdvectWithWeno5_v2, either at the advective
     time step (which becomes the next dt) or, with a constant dt, in
     ceil(dt/dt_adv) sub-steps                                          (:550-571)

The two DYREL loops are the fused solves of ht3d_v18.hip (one launch per
pseudo-iteration); the property kernels are the per-kernel entries.  The
reference builds its initial state on the CPU (SetInitialConditions_Khaled,
v18:44-128); `v18_geometry` makes a synthetic state of the same kind (a
sloping surface with air above, a fault zone, an intrusion, a geotherm and a
lithostatic-like pressure), which the caller uploads.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib, kernels as K
from .device import DeviceArray, max_abs_g, maximum_g, minimum_g
from .dimensional import Scaling, v18_scaling, v18_darcy_solve, v18_thermal_solve
from .solver import DyrelParams


@dataclass
class V18Physics:
    """Physical parameters of HydroThermal3D_v18.jl:159-186 (SI units)."""
    xmin: float = 30e3
    xmax: float = 100e3
    ymin: float = -30e3
    ymax: float = 5e3
    zmin: float = -25.0
    zmax: float = 25e3
    dT: float = 600.0            # K
    Ttop: float = 293.15         # K
    qyS: float = 37e-3           # W/m^2, basal heat flux
    dP: float = 294e6            # Pa
    Ptop: float = 1e5            # Pa
    phi_i: float = 1e-2          # porosity
    g: float = -9.81             # m/s^2
    Qt: float = 0.0              # W/m^3, heat source of intrusion cells
    T_pluton: float = 700.0 + 273.15
    delta: float = 3000.0        # m, permeability depth scale
    k_fact: float = 50.0         # fault permeability factor
    y_plateau: float = 1.5e3     # m
    C_adv: float = 5.0           # multiplier of the advective time step (v18:158)
    dt: float = 1e11             # s (v18:303)
    dt_constant: bool = True     # v18:150


def v18_geometry(nx: int, ny: int, nz: int, sc: Scaling | None = None, phys: V18Physics | None = None,
                 seed: int = 0, immersed: bool = False) -> dict:
    """Synthetic initial state on the host (Fortran arrays, scaled units):
    phases at extended centroids (phc_ex) and vertices (phv) -- 0 rock, 1 air
    above a surface rising from -17.8 km to the 1.5 km plateau, 2 a dipping
    fault zone, 3 an intrusion --, surface weights wN, a geotherm from Ttop at
    the surface to Ttop + dT at the base and a pressure from Ptop to Ptop + dP,
    both with 1 % seeded noise and air at the surface values.  With `immersed`
    the lowest air cell of each column becomes an immersed surface cell
    (phase -1) with weight in (0.5, 1)."""
    sc = sc or v18_scaling()
    ph_ = phys or V18Physics()
    rng = np.random.default_rng(seed)
    L = sc.L
    xmin, xmax, ymin, ymax = ph_.xmin / L, ph_.xmax / L, ph_.ymin / L, ph_.ymax / L
    zmin, zmax = ph_.zmin / L, ph_.zmax / L
    dx, dy, dz = (xmax - xmin) / nx, (ymax - ymin) / ny, (zmax - zmin) / nz
    xce = xmin - dx / 2 + dx * np.arange(nx + 2)
    yce = ymin - dy / 2 + dy * np.arange(ny + 2)
    zce = zmin - dz / 2 + dz * np.arange(nz + 2)
    xv, yv, zv = xmin + dx * np.arange(nx + 1), ymin + dy * np.arange(ny + 1), zmin + dz * np.arange(nz + 1)

    def surface(x):
        return np.clip(-17.8e3 / L + (x - 32e3 / L) * (17.8 / 37.4), -17.8e3 / L, ph_.y_plateau / L)

    def phases(X, Y, Z):
        p = np.zeros(X.shape)
        p[np.abs(Y - (0.5 * X - 45e3 / L)) < 2.5 * dy] = 2.0
        p[((X - 60e3 / L) / (8e3 / L)) ** 2 + ((Y + 20e3 / L) / (5e3 / L)) ** 2 +
          ((Z - 12e3 / L) / (15e3 / L)) ** 2 < 1.0] = 3.0
        p[Y > surface(X)] = 1.0
        return p

    XE, YE, ZE = np.meshgrid(xce, yce, zce, indexing="ij")
    XV, YV, ZV = np.meshgrid(xv, yv, zv, indexing="ij")
    phc = np.asfortranarray(phases(XE, YE, ZE))
    phv = np.asfortranarray(phases(XV, YV, ZV))
    wN = np.ones((nx + 2, nz + 2), order="F")
    if immersed:
        for i in range(nx + 2):
            for k in range(nz + 2):
                air = np.nonzero(phc[i, :, k] == 1.0)[0]
                if len(air) and 2 <= air[0] <= ny:
                    phc[i, air[0], k] = -1.0
                    wN[i, k] = rng.uniform(0.5, 1.0)
    Ttop, Tbot = ph_.Ttop / sc.T, (ph_.Ttop + ph_.dT) / sc.T
    Ptop, Pbot = ph_.Ptop / sc.sigma, (ph_.Ptop + ph_.dP) / sc.sigma
    ys = surface(XE)
    T = Ttop + (Ttop - Tbot) / (ys - ymin) * (YE - ys) + 0.01 * rng.standard_normal(XE.shape)
    P = Ptop + (Ptop - Pbot) / (ys - ymin) * (YE - ys) * (1.0 + 0.01 * rng.standard_normal(XE.shape))
    T[phc == 1.0] = Ttop
    P[phc == 1.0] = Ptop
    return dict(T=np.asfortranarray(T), P=np.asfortranarray(P), phc=phc, phv=phv, wN=wN,
                dx=dx, dy=dy, dz=dz, ymin=ymin, sc=sc, phys=ph_)


class V18Model:
    """Device state of the v18 model (the arrays of HydroThermal3D_v18.jl:312-340)
    and its time step.  `geometry` is a dict as returned by `v18_geometry`."""

    def __init__(self, nx: int, ny: int, nz: int, geometry: dict | None = None,
                 thermal: DyrelParams | None = None, darcy: DyrelParams | None = None):
        _lib.ensure_init()
        geo = geometry or v18_geometry(nx, ny, nz)
        self.grid = (nx, ny, nz)
        self.sc: Scaling = geo["sc"]
        self.phys: V18Physics = geo["phys"]
        self.dx, self.dy, self.dz, self.ymin = geo["dx"], geo["dy"], geo["dz"], geo["ymin"]
        self.xmin, self.zmin = self.phys.xmin / self.sc.L, self.phys.zmin / self.sc.L
        self.pT = thermal or DyrelParams(nitmax=10000, nitout=1000, tol=1e-18)     # v18:308-311
        self.pP = darcy or DyrelParams(nitmax=10000, nitout=1000, tol=1e-18)
        c, E, V = (nx, ny, nz), (nx + 2, ny + 2, nz + 2), (nx + 1, ny + 1, nz + 1)
        Z = DeviceArray.zeros
        self.a = {
            "phc_ex": DeviceArray.from_host(geo["phc"]), "phv": DeviceArray.from_host(geo["phv"]),
            "wN": DeviceArray.from_host(geo["wN"]),
            "Tc_ex": DeviceArray.from_host(geo["T"]), "Pc_ex": DeviceArray.from_host(geo["P"]),
            "Tc0": Z(c), "Pc0": Z(c), "Xc0": Z(c), "PC": Z(c), "rhoC_phirho": Z(c), "dumc": Z(c),
            "k_rhof": Z(V), "kx": Z((nx + 1, ny, nz)), "ky": Z((nx, ny + 1, nz)), "kz": Z((nx, ny, nz + 1)),
            # the smoothed hydraulic conductivity of the Darcy solve, kept for it:
            # the reference's shared vertex buffer takes the fluid density next
            "kf_v": Z(V),
            "Vx": Z((nx + 1, ny, nz)), "Vy": Z((nx, ny + 1, nz)), "Vz": Z((nx, ny, nz + 1)),
        }
        self.a["PC"].set(np.ones(c, order="F"))
        self.it = 0
        self.time = 0.0
        self.dt = self.phys.dt / self.sc.t
        self.pluton_introduced = False

    def __getitem__(self, name: str) -> DeviceArray:
        return self.a[name]

    # ------------------------------------------------------------------ phases of a step
    def _conductivity_faces(self):
        a = self.a
        K.SmoothConductivityV2C(a["dumc"], a["k_rhof"])
        K.SmoothConductivityC2V(a["k_rhof"], a["dumc"])
        K.InitConductivity(a["kx"], a["ky"], a["kz"], a["k_rhof"])

    def _hydro(self, mode: int):
        a, sc, ph = self.a, self.sc, self.phys
        K.ComputeHydroConductivity(a["k_rhof"], a["Tc_ex"], a["Pc_ex"], a["phv"], self.ymin, self.dy, ph.k_fact,
                                   ph.delta / sc.L, sc.sigma, sc.T, sc.L, sc.t, sc.eta, mode)

    def _fluid_density(self):
        a, sc = self.a, self.sc
        K.ComputeFluidDensity(a["k_rhof"], a["Tc_ex"], a["Pc_ex"], a["phv"], sc.sigma, sc.T, sc.rho)

    def thermal(self, transient: float) -> dict:
        a, sc, ph = self.a, self.sc, self.phys
        K.ComputeThermalConductivity(a["k_rhof"], a["Tc_ex"], a["phv"], ph.phi_i, sc.kt, sc.T)
        self._conductivity_faces()
        K.ComputeRhoCeffective(a["rhoC_phirho"], a["Tc_ex"], a["Pc_ex"], a["phc_ex"], ph.phi_i, sc.T, sc.sigma, sc.C,
                               sc.rho)
        return v18_thermal_solve(a["Tc_ex"], a["Tc0"], a["rhoC_phirho"], a["kx"], a["ky"], a["kz"], a["phc_ex"],
                                 a["wN"], dx=self.dx, dy=self.dy, dz=self.dz, dt=self.dt, transient=transient,
                                 qyS=ph.qyS / (sc.W / sc.L ** 2), TN=ph.Ttop / sc.T, Qt=ph.Qt / (sc.W / sc.L ** 3),
                                 norm_scale=sc.rho * sc.C * sc.T / sc.t, params=self.pT, PC=a["PC"], G=a["Xc0"],
                                 kv=a["k_rhof"])   # kx, ky, kz are its face averages

    def darcy(self, transient: float) -> dict:
        a, sc, ph = self.a, self.sc, self.phys
        self._hydro(1)
        self._conductivity_faces()
        a["kf_v"].copy_from(a["k_rhof"])   # kx, ky, kz are its face averages
        K.PhiDrhoDP(a["rhoC_phirho"], a["Tc_ex"], a["Pc_ex"], a["phc_ex"], ph.phi_i, sc.rho, sc.T, sc.sigma)
        K.PhiRhof_old(a["Xc0"], a["Tc0"], a["Pc0"], a["phc_ex"], ph.phi_i, sc.rho, sc.T, sc.sigma)
        self._fluid_density()
        return v18_darcy_solve(a["Pc_ex"], a["Tc_ex"], a["rhoC_phirho"], a["Xc0"], a["k_rhof"], a["kx"], a["ky"],
                               a["kz"], a["phc_ex"], a["wN"], dx=self.dx, dy=self.dy, dz=self.dz, dt=self.dt,
                               transient=transient, PN=ph.Ptop / sc.sigma, g=ph.g / (sc.L / sc.t ** 2),
                               phi_i=ph.phi_i, scale_rho=sc.rho, scale_T=sc.T, scale_sigma=sc.sigma,
                               norm_scale=sc.rho / sc.t, params=self.pP, PC=a["PC"], kv=a["kf_v"])

    def velocity(self) -> None:
        a, ph, sc = self.a, self.phys, self.sc
        self._hydro(2)
        self._conductivity_faces()
        self._fluid_density()
        K.ComputeDarcyFlux(a["Vx"], a["Vy"], a["Vz"], a["k_rhof"], a["kx"], a["ky"], a["kz"], a["Pc_ex"],
                           ph.g / (sc.L / sc.t ** 2), 1.0 / self.dx, 1.0 / self.dy, 1.0 / self.dz)

    def advect(self) -> int:
        """v18:550-571; returns the number of advection sub-steps."""
        a, ph, sc = self.a, self.phys, self.sc
        vmax = max(max_abs_g(a["Vx"]), max_abs_g(a["Vy"]), max_abs_g(a["Vz"]))
        dt_adv = ph.C_adv * 1.0 / 6.1 * min(self.dx, self.dy, self.dz) / vmax
        Ttop, Tbot = ph.Ttop / sc.T, (ph.Ttop + ph.dT) / sc.T
        args = (1.0 / self.dx, 1.0 / self.dy, 1.0 / self.dz, Ttop, Tbot)
        if not ph.dt_constant:
            self.dt = dt_adv
            K.AdvectWithWeno5_v2(a["Tc_ex"], None, None, a["Vx"], a["Vy"], a["Vz"], self.dt, *args)
            return 1
        n = int(math.ceil(self.dt / dt_adv))
        if n > 32767:
            raise ValueError(f"{n} advection sub-steps exceed the reference's Int16 counter")
        for _ in range(n):
            K.AdvectWithWeno5_v2(a["Tc_ex"], None, None, a["Vx"], a["Vy"], a["Vz"], self.dt / n, *args)
        return n

    def step(self) -> dict:
        """One pass of the v18 time loop (iteration `self.it`, starting at 0,
        the steady initialisation step)."""
        a, ph, sc = self.a, self.phys, self.sc
        transient = 0.0 if self.it == 0 else 1.0
        if self.it == 1 and not self.pluton_introduced:   # v18:412-414
            T = a["Tc_ex"].get()
            T[a["phc_ex"].get() == 3.0] = ph.T_pluton / sc.T
            a["Tc_ex"].set(T)
            self.pluton_introduced = True
        K.Cpy_inn_to_all(a["Tc0"], a["Tc_ex"])
        K.Cpy_inn_to_all(a["Pc0"], a["Pc_ex"])
        rT = self.thermal(transient)
        rP = self.darcy(transient)
        self.velocity()
        n_adv = 0
        dt_used = self.dt
        if self.it > 0:
            self.time += self.dt
            n_adv = self.advect()
        out = dict(it=self.it, thermal=rT, darcy=rP, dt=dt_used, dt_next=self.dt, n_adv=n_adv, time=self.time)
        out.update(self.extrema())
        self.it += 1
        return out

    def extrema(self) -> dict:
        """min/max of Tc_ex (K), Pc_ex (Pa) and Vy (m/s), the diagnostics the v18
        driver prints after each step (v18:575-577), as device reductions (8 bytes
        read back each) instead of host copies."""
        a, sc = self.a, self.sc
        return {"T_min": minimum_g(a["Tc_ex"]) * sc.T, "T_max": maximum_g(a["Tc_ex"]) * sc.T,
                "P_min": minimum_g(a["Pc_ex"]) * sc.sigma, "P_max": maximum_g(a["Pc_ex"]) * sc.sigma,
                "Vy_min": minimum_g(a["Vy"]) * sc.V, "Vy_max": maximum_g(a["Vy"]) * sc.V}

    # ------------------------------------------------------------------ output and restart
    def output_fields(self) -> dict:
        """The VTK fields of HydroThermal3D_v18.jl:671-686, computed on the
        device by the kernels of kernels_HT3D_postprocessing.jl in the same
        order through one scratch array (so, as there, HeatCapacity is NaN in
        air, left by Peclet).  Velocity's second and third components start
        from zero arrays where the reference reuses the solver's residual
        arrays (they keep those values in air)."""
        a, sc, ph = self.a, self.sc, self.phys
        c = self.grid
        dumc = a["dumc"]
        X, Y, Z = DeviceArray.zeros(c), DeviceArray.zeros(c), DeviceArray.zeros(c)
        out = {}
        K.Phase(dumc, a["phc_ex"]); out["Phase"] = dumc.get()
        K.Pressure(dumc, a["Pc_ex"], a["phc_ex"], sc.sigma); out["Pressure"] = dumc.get()
        K.Temperature(dumc, a["Tc_ex"], a["phc_ex"], sc.T); out["Temperature"] = dumc.get()
        K.Velocity(X, Y, Z, a["Vx"], a["Vy"], a["Vz"], a["phc_ex"], sc.V)
        out["Velocity"] = (X.get(), Y.get(), Z.get())
        K.EffectiveThermalConductivity(dumc, a["Tc_ex"], ph.phi_i, a["phc_ex"], sc.T); out["kThermal"] = dumc.get()
        K.Permeability(dumc, ph.k_fact, self.ymin, self.dy, a["phc_ex"], ph.delta / sc.L, sc.L)
        out["Perm"] = dumc.get()
        K.FluidDensity(dumc, a["Tc_ex"], a["Pc_ex"], a["phc_ex"], sc.T, sc.sigma); out["Density"] = dumc.get()
        K.Viscosity(dumc, a["Tc_ex"], a["phc_ex"], sc.T); out["Viscosity"] = dumc.get()
        K.Peclet(dumc, a["Tc_ex"], a["Pc_ex"], a["Vx"], a["Vy"], a["Vz"], a["phc_ex"], ph.phi_i, sc.T, sc.sigma, sc.V,
                 sc.L)
        out["Peclet"] = dumc.get()
        K.EffectiveHeatCapacity(dumc, a["Tc_ex"], a["Pc_ex"], a["phc_ex"], ph.phi_i, sc.T, sc.sigma, sc.C, sc.rho)
        out["HeatCapacity"] = dumc.get()
        return out

    def cell_centres(self):
        nx, ny, nz = self.grid
        return (self.xmin + self.dx / 2 + self.dx * np.arange(nx), self.ymin + self.dy / 2 + self.dy * np.arange(ny),
                self.zmin + self.dz / 2 + self.dz * np.arange(nz))

    def save_vtk(self, name) -> "Path":
        """HT3DOutput%05d.vtr of v18:663-687 (grid points at the cell centres)."""
        from .vtkio import write_vtr
        return write_vtr(name, *self.cell_centres(), self.output_fields())

    def save_breakpoint(self, name) -> "Path":
        """Breakpoint%05d (v18:579-588): Tc_ex, Pc_ex, time and dt."""
        from .vtkio import save_breakpoint
        return save_breakpoint(name, self.a["Tc_ex"].get(), self.a["Pc_ex"].get(), self.time, self.dt)

    def restart(self, path, it: int) -> None:
        """Restart from a breakpoint at time step `it` (v18:384-395)."""
        from .vtkio import load_breakpoint
        b = load_breakpoint(path)
        self.a["Tc_ex"].set(b["Tc_ex"])
        self.a["Pc_ex"].set(b["Pc_ex"])
        self.time, self.dt = b["time"], b["dt"]
        self.it = it + 1
        self.pluton_introduced = self.it > 1
