"""Shared helpers of the test suite (staggered shapes, random inputs)."""
import numpy as np


def shapes(nx, ny, nz):
    return {
        "c": (nx, ny, nz), "E": (nx + 2, ny + 2, nz + 2), "X6": (nx + 6, ny + 6, nz + 6),
        "V": (nx + 1, ny + 1, nz + 1), "Fx": (nx + 1, ny, nz), "Fy": (nx, ny + 1, nz), "Fz": (nx, ny, nz + 1),
    }


def rand(shape, rng, lo=-1.0, hi=1.0):
    return np.asfortranarray(rng.uniform(lo, hi, size=shape))


def positive(shape, rng):
    return np.asfortranarray(10.0 ** rng.uniform(-1.0, 1.0, size=shape))


def face_ghost_mask(nx, ny, nz):
    """Interior plus face ghosts of an E array (edges/corners excluded)."""
    m = np.zeros((nx + 2, ny + 2, nz + 2), dtype=bool)
    m[1:-1, 1:-1, 1:-1] = True
    m[[0, -1], 1:-1, 1:-1] = True
    m[1:-1, [0, -1], 1:-1] = True
    m[1:-1, 1:-1, [0, -1]] = True
    return m


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / den)


def dim_state(nx, ny, nz, rng, air_from=None, immersed=False, TN=None, PN=None):
    """Synthetic inputs of the dimensional (v18) kernels in the model's scaled
    units: T in [293, 1000] K and P in [1e5, 3e8] Pa divided by their scales,
    vertex and centroid phases in {0, 2, 3} below the row `air_from` and air (1)
    from it up (ghost rows included), air cells already at the surface values
    (as the v18 initial conditions leave them), and optionally one immersed
    surface cell (phase -1) per column, directly below the air, away from the
    ghost layers.  Returns a dict of Fortran-ordered arrays and scalars."""
    from hydrothermal3d_amd.dimensional import v18_scaling
    sc = v18_scaling()
    sh = shapes(nx, ny, nz)
    T = np.asfortranarray(rng.uniform(293.0, 1000.0, sh["E"]) / sc.T)
    P = np.asfortranarray(rng.uniform(1e5, 3e8, sh["E"]) / sc.sigma)
    phc = np.asfortranarray(rng.choice([0.0, 2.0, 3.0], size=sh["E"]))
    phv = np.asfortranarray(rng.choice([0.0, 2.0, 3.0], size=sh["V"]))
    TN = 293.15 / sc.T if TN is None else TN
    PN = 1e5 / sc.sigma if PN is None else PN
    wN = np.asfortranarray(rng.uniform(0.3, 1.0, size=(nx + 2, nz + 2)))
    if air_from is not None:
        phc[:, air_from:, :] = 1.0
        phv[:, air_from:, :] = 1.0
        T[phc == 1.0] = TN
        P[phc == 1.0] = PN
        if immersed:
            cols = [(i, k) for i in range(2, nx) for k in range(2, nz) if rng.uniform() < 0.6]
            for i, k in cols:
                phc[i, air_from - 1, k] = -1.0
    return dict(sc=sc, T=T, P=P, phc=phc, phv=phv, TN=TN, PN=PN, wN=wN, sh=sh)


def v18_problem(nx, ny, nz, seed=0, immersed=False, Qt=None):
    """A small dimensional (v18) problem in the reference's scaled units, with
    the property set-up of Dimensional/HydroThermal3D_v18.jl:426-430 (heat) and
    :484-503 (Darcy) done by the CPU oracle: a domain 70 x 35 x 50 km with a
    sloping surface (air above), a fault band (phase 2), an intrusion (phase 3),
    a geotherm and a hydrostatic-like pressure, optionally an immersed-surface
    layer (phase -1, the lowest air cell of each column, weight in (0.2, 1)).
    Returns a dict of Fortran arrays and scalars for both solves."""
    from oracle import oracle as orc
    from hydrothermal3d_amd.dimensional import v18_scaling
    sc = v18_scaling()
    rng = np.random.default_rng(seed)
    sh = shapes(nx, ny, nz)
    xmin, xmax, ymin, ymax, zmin, zmax = 30e3 / sc.L, 100e3 / sc.L, -30e3 / sc.L, 5e3 / sc.L, -25.0 / sc.L, 25e3 / sc.L
    dx, dy, dz = (xmax - xmin) / nx, (ymax - ymin) / ny, (zmax - zmin) / nz
    xce = xmin - dx / 2 + dx * np.arange(nx + 2)
    yce = ymin - dy / 2 + dy * np.arange(ny + 2)
    zce = zmin - dz / 2 + dz * np.arange(nz + 2)
    xv, yv, zv = xmin + dx * np.arange(nx + 1), ymin + dy * np.arange(ny + 1), zmin + dz * np.arange(nz + 1)

    def surface(x):
        return np.clip(-17.8e3 / sc.L + (x - 32e3 / sc.L) * (17.8 / 37.4), -17.8e3 / sc.L, 1.5e3 / sc.L)

    def phases(X, Y, Z):
        ph = np.zeros(X.shape)
        ph[np.abs(Y - (0.5 * X - 45e3 / sc.L)) < 2.5 * dy] = 2.0
        ph[((X - 60e3 / sc.L) / (8e3 / sc.L)) ** 2 + ((Y + 20e3 / sc.L) / (5e3 / sc.L)) ** 2 +
           ((Z - 12e3 / sc.L) / (15e3 / sc.L)) ** 2 < 1.0] = 3.0
        ph[Y > surface(X)] = 1.0
        return ph

    XE, YE, ZE = np.meshgrid(xce, yce, zce, indexing="ij")
    XV, YV, ZV = np.meshgrid(xv, yv, zv, indexing="ij")
    phc = np.asfortranarray(phases(XE, YE, ZE))
    phv = np.asfortranarray(phases(XV, YV, ZV))
    wN = np.ones((nx + 2, nz + 2), order="F")
    if immersed:
        for i in range(nx + 2):
            for k in range(nz + 2):
                air = np.nonzero(phc[i, :, k] == 1.0)[0]
                if len(air) and 2 <= air[0] <= ny:       # interior cell, the cell below is rock
                    phc[i, air[0], k] = -1.0
                    wN[i, k] = rng.uniform(0.2, 1.0)
    Ttop, Tbot = 293.15 / sc.T, (293.15 + 600.0) / sc.T
    Ptop, Pbot = 1e5 / sc.sigma, (1e5 + 294e6) / sc.sigma
    ys = surface(XE)
    T = Ttop + (Ttop - Tbot) / (ys - ymin) * (YE - ys) + 0.01 * rng.standard_normal(XE.shape)
    P = Ptop + (Ptop - Pbot) / (ys - ymin) * (YE - ys) * (1.0 + 0.01 * rng.standard_normal(XE.shape))
    T[phc == 1.0] = Ttop
    P[phc == 1.0] = Ptop
    T, P = np.asfortranarray(T), np.asfortranarray(P)
    phi_i, delta, k_fact = 1e-2, 3000.0 / sc.L, 50.0
    z = lambda s: np.zeros(sh[s], order="F")
    # heat set-up (v18:426-430)
    ktv, dumc = z("V"), z("c")
    orc.ComputeThermalConductivity(ktv, T, phv, phi_i, sc.kt, sc.T)
    orc.SmoothConductivityV2C(dumc, ktv)
    orc.SmoothConductivityC2V(ktv, dumc)
    kT = [z("Fx"), z("Fy"), z("Fz")]
    orc.InitConductivity(*kT, ktv)
    rhoC = z("c")
    orc.ComputeRhoCeffective(rhoC, T, P, phc, phi_i, sc.T, sc.sigma, sc.C, sc.rho)
    T0 = np.asfortranarray(T[1:-1, 1:-1, 1:-1])
    # Darcy set-up (v18:484-503)
    kfv = z("V")
    orc.ComputeHydroConductivity(kfv, T, P, phv, ymin, dy, k_fact, delta, sc.sigma, sc.T, sc.L, sc.t, sc.eta, 1)
    orc.SmoothConductivityV2C(dumc, kfv)
    orc.SmoothConductivityC2V(kfv, dumc)
    kP = [z("Fx"), z("Fy"), z("Fz")]
    orc.InitConductivity(*kP, kfv)
    phirho = z("c")
    orc.PhiDrhoDP(phirho, T, P, phc, phi_i, sc.rho, sc.T, sc.sigma)
    P0 = np.asfortranarray(P[1:-1, 1:-1, 1:-1])
    phirhof0 = z("c")
    orc.PhiRhof_old(phirhof0, T0, P0, phc, phi_i, sc.rho, sc.T, sc.sigma)
    rhofv = z("V")
    orc.ComputeFluidDensity(rhofv, T, P, phv, sc.sigma, sc.T, sc.rho)
    dt = 1e11 / sc.t
    heat = dict(dx=dx, dy=dy, dz=dz, dt=dt, transient=1.0, qyS=37e-3 / (sc.W / sc.L ** 2), TN=Ttop,
                Qt=(3.4e-4 / (sc.W / sc.L ** 3)) if Qt is None else Qt, norm_scale=sc.rho * sc.C * sc.T / sc.t)
    darcy = dict(dx=dx, dy=dy, dz=dz, dt=dt, transient=1.0, PN=Ptop, g=-9.81 / (sc.L / sc.t ** 2), phi_i=phi_i,
                 scale_rho=sc.rho, scale_T=sc.T, scale_sigma=sc.sigma, norm_scale=sc.rho / sc.t)
    return dict(T=T, P=P, T0=T0, rhoC=rhoC, kT=kT, phirho=phirho, phirhof0=phirhof0, rhofv=rhofv, kP=kP, phc=phc,
                wN=wN, heat=heat, darcy=darcy, sc=sc, ktv=ktv, kfv=kfv)


class V18OracleModel:
    """The v18 time step of HydroThermal3D_v18.jl:398-571 with the CPU oracle's
    kernels and unfused DYREL loops, on the same host state and in the same
    order as hydrothermal3d_amd.v18.V18Model (the oracle for that driver)."""

    def __init__(self, nx, ny, nz, geo, pT, pP, ulp_seed=None):
        """ulp_seed: if given, every permeability from ComputeHydroConductivity
        (whose exp() and 10^x may round differently on another machine) is moved
        by one ulp at seeded random cells, to measure the result's sensitivity."""
        from oracle import oracle as orc
        self.ulp_rng = None if ulp_seed is None else np.random.default_rng(ulp_seed)
        self.orc, self.geo, self.sc, self.ph = orc, geo, geo["sc"], geo["phys"]
        self.pT, self.pP = pT, pP
        sh = shapes(nx, ny, nz)
        z = lambda s: np.zeros(sh[s], order="F")
        self.T, self.P = geo["T"].copy(order="F"), geo["P"].copy(order="F")
        self.phc, self.phv, self.wN = geo["phc"], geo["phv"], geo["wN"]
        self.Tc0, self.Pc0, self.Xc0, self.rho, self.dumc = z("c"), z("c"), z("c"), z("c"), z("c")
        self.k = z("V")
        self.kx, self.ky, self.kz = z("Fx"), z("Fy"), z("Fz")
        self.Vx, self.Vy, self.Vz = z("Fx"), z("Fy"), z("Fz")
        self.it, self.time, self.dt = 0, 0.0, self.ph.dt / self.sc.t

    def _hydro(self, *args):
        self.orc.ComputeHydroConductivity(self.k, *args)
        if self.ulp_rng is not None:
            sign = self.ulp_rng.choice([-1.0, 0.0, 1.0], size=self.k.shape)
            self.k[...] = np.where(sign > 0, np.nextafter(self.k, np.inf),
                                   np.where(sign < 0, np.nextafter(self.k, -np.inf), self.k))

    def _faces(self):
        o = self.orc
        o.SmoothConductivityV2C(self.dumc, self.k)
        o.SmoothConductivityC2V(self.k, self.dumc)
        o.InitConductivity(self.kx, self.ky, self.kz, self.k)

    def step(self):
        import math
        o, sc, ph, g = self.orc, self.sc, self.ph, self.geo
        transient = 0.0 if self.it == 0 else 1.0
        if self.it == 1:
            self.T[self.phc == 3.0] = ph.T_pluton / sc.T
        self.Tc0[...] = self.T[1:-1, 1:-1, 1:-1]
        self.Pc0[...] = self.P[1:-1, 1:-1, 1:-1]
        h = dict(dx=g["dx"], dy=g["dy"], dz=g["dz"], dt=self.dt, transient=transient)
        # heat
        o.ComputeThermalConductivity(self.k, self.T, self.phv, ph.phi_i, sc.kt, sc.T)
        self._faces()
        o.ComputeRhoCeffective(self.rho, self.T, self.P, self.phc, ph.phi_i, sc.T, sc.sigma, sc.C, sc.rho)
        rT, _ = o.v18_thermal_solve(self.T, self.Tc0, self.rho, self.kx, self.ky, self.kz, self.phc, self.wN, **h,
                                    qyS=ph.qyS / (sc.W / sc.L ** 2), TN=ph.Ttop / sc.T, Qt=ph.Qt / (sc.W / sc.L ** 3),
                                    norm_scale=sc.rho * sc.C * sc.T / sc.t, p=self.pT)
        # Darcy
        args = (self.T, self.P, self.phv, g["ymin"], g["dy"], ph.k_fact, ph.delta / sc.L, sc.sigma, sc.T, sc.L, sc.t,
                sc.eta)
        self._hydro(*args, 1)
        self._faces()
        o.PhiDrhoDP(self.rho, self.T, self.P, self.phc, ph.phi_i, sc.rho, sc.T, sc.sigma)
        o.PhiRhof_old(self.Xc0, self.Tc0, self.Pc0, self.phc, ph.phi_i, sc.rho, sc.T, sc.sigma)
        o.ComputeFluidDensity(self.k, self.T, self.P, self.phv, sc.sigma, sc.T, sc.rho)
        grav = ph.g / (sc.L / sc.t ** 2)
        rP, _ = o.v18_darcy_solve(self.P, self.T, self.rho, self.Xc0, self.k, self.kx, self.ky, self.kz, self.phc,
                                  self.wN, **h, PN=ph.Ptop / sc.sigma, g=grav, phi_i=ph.phi_i, scale_rho=sc.rho,
                                  scale_T=sc.T, scale_sigma=sc.sigma, norm_scale=sc.rho / sc.t, p=self.pP)
        # velocity
        self._hydro(*args, 2)
        self._faces()
        o.ComputeFluidDensity(self.k, self.T, self.P, self.phv, sc.sigma, sc.T, sc.rho)
        o.ComputeDarcyFlux(self.Vx, self.Vy, self.Vz, self.k, self.kx, self.ky, self.kz, self.P, grav,
                           1 / g["dx"], 1 / g["dy"], 1 / g["dz"])
        n_adv = 0
        if self.it > 0:
            self.time += self.dt
            vmax = max(np.max(np.abs(self.Vx)), np.max(np.abs(self.Vy)), np.max(np.abs(self.Vz)))
            dt_adv = ph.C_adv * 1.0 / 6.1 * min(g["dx"], g["dy"], g["dz"]) / vmax
            n_adv = int(math.ceil(self.dt / dt_adv))
            Told, T6 = np.zeros(self.Tc0.shape, order="F"), np.zeros(tuple(s + 4 for s in self.T.shape), order="F")
            for _ in range(n_adv):
                o.AdvectWithWeno5_v2(self.T, T6, Told, self.Vx, self.Vy, self.Vz, self.dt / n_adv, 1 / g["dx"],
                                     1 / g["dy"], 1 / g["dz"], ph.Ttop / sc.T, (ph.Ttop + ph.dT) / sc.T)
        self.it += 1
        return dict(thermal=rT, darcy=rP, n_adv=n_adv)
