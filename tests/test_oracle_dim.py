"""Pins of the oracle's dimensional (v18) property kernels and closures
(oracle/ht3d_oracle_dim.c) -- CPU only.

* the pressure derivative of the fluid-density polynomial against the output
  stored in the reference notebook (notebooks/Gershgorin_TH3D.ipynb), both as
  the closed-form derivative closure and as a finite difference of rho_f;
* the literal fourth power against exact rational arithmetic;
* the water-viscosity closure against tabulated values;
* every kernel against the independent NumPy restatement;
* closed forms (uniform fields, the permeability profile, the fault factor,
  air cells, the immersed-surface interpolation identity);
* the ghost fills against a literal one-thread execution of the reference
  kernel body.
"""
import json
import re
from fractions import Fraction
from pathlib import Path

import numpy as np
import pytest

from helpers import dim_state, shapes

GOLD = Path(__file__).resolve().parent / "golden" / "notebook_outputs.json"


def _julia_scalar_fn(expr: str):
    """The notebook's printed Julia expression as a Python function of (T_C, p),
    integer powers written out as products (Julia's literal powers)."""
    def powrep(m):
        v, n = m.group(1), int(m.group(2))
        return "(" + "*".join([v] * n) + ")"
    py = re.sub(r"(\w+) \.\^ (\d)", powrep, expr).replace(".*", "*")
    return eval("lambda T_C, p: " + py)  # noqa: S307 (fixture text from the repository)


@pytest.fixture(scope="module")
def npd():
    from oracle import ht3d_numpy_dim
    return ht3d_numpy_dim


def _samples(rng, n=400):
    return rng.uniform(0.0, 700.0, n), rng.uniform(1e5, 3e8, n)


def test_drhodp_matches_notebook_output(orc):
    d_nb = _julia_scalar_fn(json.loads(GOLD.read_text())["drhof_dp"])
    rng = np.random.default_rng(1)
    Tc, p = _samples(rng)
    got = np.array([orc.eos("drhodP_C", a, b) for a, b in zip(Tc, p)])
    want = np.array([d_nb(a, b) for a, b in zip(Tc, p)])
    # same terms and order: equal up to the last bit of the largest term
    assert np.max(np.abs(got - want)) <= 4e-16 * np.max(np.abs(want))


def test_rhof_pressure_slope_matches_notebook_derivative(orc):
    """Central differences of the oracle's rho_f(T, p) in p reproduce the
    notebook's derivative: pins every pressure-dependent coefficient of rho_f."""
    d_nb = _julia_scalar_fn(json.loads(GOLD.read_text())["drhof_dp"])
    rng = np.random.default_rng(2)
    for Tc, p in zip(*_samples(rng, 200)):
        h = 2e4
        fd = (orc.eos("rhof_C", Tc, p + h) - orc.eos("rhof_C", Tc, p - h)) / (2 * h)
        assert abs(fd - d_nb(Tc, p)) <= 1e-6 * abs(d_nb(Tc, p)) + 1e-14


def test_rhof_temperature_terms_match_numpy(orc, npd):
    """The pressure-independent terms of rho_f (not reached by the derivative
    pin) are checked against the independent restatement, bit for bit."""
    rng = np.random.default_rng(3)
    Tc, p = _samples(rng)
    p[:50] = 0.0
    got = np.array([orc.eos("rhof_C", a, b) for a, b in zip(Tc, p)])
    assert np.array_equal(got, npd.rhof_C(Tc, p))
    assert np.array_equal(np.array([orc.eos("drhodP_C", a, b) for a, b in zip(Tc, p)]), npd.drhodP_C(Tc, p))
    for name in ("kTs", "kTf", "Cf", "Cs", "rhos"):
        T = rng.uniform(250.0, 1200.0, 50)
        assert np.array_equal(np.array([orc.eos(name, t) for t in T]), getattr(npd, name)(T)), name


def test_pow4_is_the_rounded_fourth_power(orc, npd):
    rng = np.random.default_rng(4)
    xs = np.concatenate([rng.uniform(1e5, 3e8, 300), rng.uniform(0.0, 700.0, 300), [3.0, 1e8, 0.1]])
    exact = 0
    for x in xs:
        got = orc.eos("pow4", x)
        want = float(Fraction(x) ** 4)
        assert abs(got - want) <= np.spacing(want)
        exact += got == want
        assert npd.pow4(np.float64(x)) == got
    assert exact >= len(xs) - 3


@pytest.mark.parametrize("T_K,mu", [(293.15, 1.0016e-3), (323.15, 5.47e-4), (373.15, 2.82e-4)])
def test_water_viscosity(orc, T_K, mu):
    """mu_f(T) against tabulated liquid-water viscosities (20, 50, 100 C); the
    closure is a Vogel-type fit, accurate to about 2 % over liquid water."""
    assert abs(orc.eos("muf", T_K) / mu - 1.0) < 0.02


def test_phi_and_permeability_law(orc):
    assert orc.eos("phi", 2.0, 0.01) == 0.03 and orc.eos("phi", 0.0, 0.01) == 0.01
    assert orc.eos("phi", 3.0, 0.01) == 0.01
    assert orc.eos("kF", 0.0, 3000.0) == 5e-16
    assert abs(orc.eos("kF", -3000.0, 3000.0) - 5e-16 / np.e) < 1e-30


GRIDS = [(6, 5, 4), (9, 7, 8)]


@pytest.mark.parametrize("grid", GRIDS, ids=lambda g: "x".join(map(str, g)))
def test_property_kernels_match_numpy(orc, npd, grid):
    rng = np.random.default_rng(sum(grid))
    d = dim_state(*grid, rng, air_from=grid[1] - 1)
    sc, sh = d["sc"], d["sh"]
    z = lambda s: np.zeros(sh[s], order="F")
    # thermal conductivity at vertices
    ktv = z("V")
    orc.ComputeThermalConductivity(ktv, d["T"], d["phv"], 0.01, sc.kt, sc.T)
    assert np.array_equal(ktv, npd.ComputeThermalConductivity(d["T"], d["phv"], 0.01, sc.kt, sc.T))
    # smoothing vertex -> cell -> vertex
    ktc = z("c")
    orc.SmoothConductivityV2C(ktc, ktv)
    assert np.array_equal(ktc, npd.SmoothConductivityV2C(ktv))
    ktv2 = ktv.copy(order="F")
    orc.SmoothConductivityC2V(ktv2, ktc)
    assert np.array_equal(ktv2, npd.SmoothConductivityC2V(ktv, ktc))
    # hydraulic conductivity, three modes (exp differs by an ulp between libm and NumPy)
    for mode in (1, 2, 3):
        kfv = np.asfortranarray(rng.uniform(1.0, 2.0, sh["V"]))
        want = npd.ComputeHydroConductivity(kfv, d["T"], d["P"], d["phv"], -30e3 / sc.L, 35e3 / sc.L / grid[1],
                                            50.0, 3000.0 / sc.L, sc.sigma, sc.T, sc.L, sc.t, sc.eta, mode)
        orc.ComputeHydroConductivity(kfv, d["T"], d["P"], d["phv"], -30e3 / sc.L, 35e3 / sc.L / grid[1], 50.0,
                                     3000.0 / sc.L, sc.sigma, sc.T, sc.L, sc.t, sc.eta, mode)
        np.testing.assert_allclose(kfv, want, rtol=1e-15, atol=0)
    # fluid density at vertices
    rhofv = z("V")
    orc.ComputeFluidDensity(rhofv, d["T"], d["P"], d["phv"], sc.sigma, sc.T, sc.rho)
    assert np.array_equal(rhofv, npd.ComputeFluidDensity(d["T"], d["P"], d["phv"], sc.sigma, sc.T, sc.rho))
    # effective rho*C (air cells untouched)
    rc0 = np.asfortranarray(rng.uniform(-1, 1, sh["c"]))
    rc = rc0.copy(order="F")
    orc.ComputeRhoCeffective(rc, d["T"], d["P"], d["phc"], 0.01, sc.T, sc.sigma, sc.C, sc.rho)
    assert np.array_equal(rc, npd.ComputeRhoCeffective(rc0, d["T"], d["P"], d["phc"], 0.01, sc.T, sc.sigma, sc.C,
                                                       sc.rho))
    # phi rho_f (extended and cell inputs) and phi drho/dP
    for fn, T, P, deriv in (("PhiRhof", d["T"], d["P"], False),
                            ("PhiRhof_old", np.asfortranarray(d["T"][1:-1, 1:-1, 1:-1]),
                             np.asfortranarray(d["P"][1:-1, 1:-1, 1:-1]), False),
                            ("PhiDrhoDP", d["T"], d["P"], True)):
        out0 = np.asfortranarray(rng.uniform(-1, 1, sh["c"]))
        out = out0.copy(order="F")
        getattr(orc, fn)(out, T, P, d["phc"], 0.01, sc.rho, sc.T, sc.sigma)
        assert np.array_equal(out, npd.PhiRhof(out0, T, P, d["phc"], 0.01, sc.rho, sc.T, sc.sigma, deriv)), fn
    # dimensional Init_vel
    q = [np.asfortranarray(rng.uniform(-1, 1, sh[s])) for s in ("Fx", "Fy", "Fz")]
    rho = np.asfortranarray(rng.uniform(0.5, 2.0, sh["V"]))
    V = [z(s) for s in ("Fx", "Fy", "Fz")]
    orc.dim_Init_vel(*V, *q, rho)
    for a, b in zip(V, npd.dim_Init_vel(*q, rho)):
        assert np.array_equal(a, b)


def test_property_closed_forms(orc):
    nx, ny, nz = 5, 6, 4
    sh = shapes(nx, ny, nz)
    from hydrothermal3d_amd.dimensional import v18_scaling
    sc = v18_scaling()
    T_K, P_Pa, phi0 = 600.0, 5e7, 0.02
    T = np.full(sh["E"], T_K / sc.T, order="F")
    P = np.full(sh["E"], P_Pa / sc.sigma, order="F")
    phv = np.zeros(sh["V"], order="F")
    phv[2, :, :] = 2.0
    phv[:, -1, :] = 1.0
    # uniform T: conductivity is the phase mixture of the closures at T (up to the
    # rounding of the eight-point mean)
    ktv = np.zeros(sh["V"], order="F")
    orc.ComputeThermalConductivity(ktv, T, phv, phi0, sc.kt, sc.T)
    Tv = ktv * 0 + T_K
    f = np.where(phv == 2.0, 3 * phi0, phi0)
    want = ((1 - f) * (3.1138 - 0.0023 * Tv) + f * (-0.069 + 0.0012 * Tv)) / sc.kt
    np.testing.assert_allclose(ktv, want, rtol=1e-14)
    # a constant survives vertex -> cell -> vertex smoothing
    c = np.zeros(sh["c"], order="F")
    k1 = np.full(sh["V"], 0.3, order="F")
    orc.SmoothConductivityV2C(c, k1)
    orc.SmoothConductivityC2V(k1, c)
    np.testing.assert_allclose(k1, 0.3, rtol=2e-16)
    # a field linear in (x, y, z) is averaged to its cell-centre value
    xv, yv, zv = np.meshgrid(np.arange(nx + 1.0), np.arange(ny + 1.0), np.arange(nz + 1.0), indexing="ij")
    lin = np.asfortranarray(1.0 + 2.0 * xv - 3.0 * yv + 0.5 * zv)
    orc.SmoothConductivityV2C(c, lin)
    xc, yc, zc = np.meshgrid(np.arange(nx) + 0.5, np.arange(ny) + 0.5, np.arange(nz) + 0.5, indexing="ij")
    np.testing.assert_allclose(c, 1.0 + 2.0 * xc - 3.0 * yc + 0.5 * zc, rtol=1e-14)
    # permeability law (mode 3): 5e-16 exp(y/delta) / L^2, times k_fact on faults
    ymin, dy, delta, kfact = -30e3 / sc.L, 35e3 / sc.L / ny, 3000.0 / sc.L, 50.0
    kfv = np.zeros(sh["V"], order="F")
    orc.ComputeHydroConductivity(kfv, T, P, phv, ymin, dy, kfact, delta, sc.sigma, sc.T, sc.L, sc.t, sc.eta, 3)
    y = ymin + np.arange(ny + 1) * dy
    prof = 5e-16 * np.exp(y * sc.L / (delta * sc.L)) / (sc.L * sc.L)
    want = np.broadcast_to(prof[None, :, None], sh["V"]) * np.where(phv == 2.0, kfact, 1.0)
    np.testing.assert_allclose(kfv, want, rtol=1e-15)
    # mode 1 / mode 2 ratio is rho_f(T, P) * scale_L^2 / (scale_eta * scale_t)
    k1v, k2v = np.zeros(sh["V"], order="F"), np.zeros(sh["V"], order="F")
    args = (T, P, phv, ymin, dy, kfact, delta, sc.sigma, sc.T, sc.L, sc.t, sc.eta)
    orc.ComputeHydroConductivity(k1v, *args, 1)
    orc.ComputeHydroConductivity(k2v, *args, 2)
    rho = orc.eos("rhof_C", T_K - 273.15, P_Pa)
    np.testing.assert_allclose(k1v / k2v, rho * sc.L ** 2 / sc.eta / sc.t, rtol=1e-13)
    # fluid density: rho_f / scale_rho, zero in air
    rf = np.zeros(sh["V"], order="F")
    orc.ComputeFluidDensity(rf, T, P, phv, sc.sigma, sc.T, sc.rho)
    np.testing.assert_allclose(rf[phv != 1.0], rho / sc.rho, rtol=1e-14)
    assert np.all(rf[phv == 1.0] == 0.0)


@pytest.mark.parametrize("immersed", [False, True], ids=["flat", "immersed"])
@pytest.mark.parametrize("grid", GRIDS, ids=lambda g: "x".join(map(str, g)))
def test_ghost_fills_match_serial_reference(orc, npd, grid, immersed):
    nx, ny, nz = grid
    rng = np.random.default_rng(7 + nx)
    d = dim_state(nx, ny, nz, rng, air_from=ny - 1, immersed=immersed)
    sh = d["sh"]
    ky = np.asfortranarray(rng.uniform(0.5, 2.0, sh["Fy"]))
    qyS, _dy = 37e-3 / d["sc"].W * d["sc"].L ** 2, ny / 3.5
    # temperature
    T = d["T"].copy(order="F")
    orc.dim_SetTemperatureBCs(T, ky, d["wN"], d["phc"], qyS, _dy, 1.05, 0.0, d["TN"], *([0.0] * 10))
    Ts = npd.SetTemperatureBCs_serial(d["T"], ky, d["wN"], d["phc"], qyS, _dy, d["TN"])
    m = np.ones(sh["E"], dtype=bool)
    m[[0, -1], :, -1] = False          # x-face/z-face shared lines: thread-order dependent in the reference
    assert np.array_equal(T[m], Ts[m])
    # every ghost the hot path reads was written
    assert np.all(T[1:-1, 0, 1:-1] != d["T"][1:-1, 0, 1:-1])
    # the immersed cell interpolates the surface value with the cell below
    if immersed:
        idx = np.argwhere(d["phc"] == -1.0)
        assert len(idx) > 0
        for i, j, k in idx:
            w = d["wN"][i, k]
            assert abs(w * T[i, j, k] + (1 - w) * T[i, j - 1, k] - d["TN"]) < 1e-14 * abs(d["TN"]) / w
    # pressure
    P = d["P"].copy(order="F")
    orc.dim_SetPressureBCs(P, d["wN"], d["phc"], 0.0, d["PN"], *([0.0] * 8))
    Ps = npd.SetPressureBCs_serial(d["P"], d["wN"], d["phc"], d["PN"])
    assert np.array_equal(P, Ps)
    assert np.array_equal(P[1:-1, 0, 1:-1], P[1:-1, 1, 1:-1])


# ------------------------------------------------------------ v18 DYREL solves
def _v18(grid, immersed=False, seed=0, dt_s=1e15, Qt=None):
    from helpers import v18_problem
    d = v18_problem(*grid, seed=seed, immersed=immersed, Qt=Qt)
    d["heat"].update(dt=dt_s / d["sc"].t, norm_scale=1e30)    # relative convergence only
    d["darcy"].update(dt=dt_s / d["sc"].t, norm_scale=1e30)
    return d


def test_v18_residuals_match_numpy(orc, npd):
    d = _v18((9, 8, 5), immersed=True)
    h, dd = d["heat"], d["darcy"]
    sh = shapes(9, 8, 5)
    rng = np.random.default_rng(5)
    _d = (1 / h["dx"], 1 / h["dy"], 1 / h["dz"])
    q = [np.zeros(sh[s], order="F") for s in ("Fx", "Fy", "Fz")]
    orc.ComputeFlux(*q, *d["kT"], d["T"], *_d)
    qn = npd.ComputeFlux(*d["kT"], d["T"], *_d)
    for a, b in zip(q, qn):
        assert np.array_equal(a, b)
    PC = np.asfortranarray(rng.uniform(0.5, 2.0, sh["c"]))
    F = np.zeros(sh["c"], order="F")
    orc.ResidualTemperatureLinearised(F, d["T"], d["T0"], d["rhoC"], d["phc"], PC, *q, h["Qt"], 1.0, 1 / h["dt"], *_d)
    assert np.array_equal(F, npd.ResidualTemperatureLinearised(d["T"], d["T0"], d["rhoC"], d["phc"], PC, *qn,
                                                               h["Qt"], 1.0, 1 / h["dt"], *_d))
    orc.ComputeDarcyFlux(*q, d["rhofv"], *d["kP"], d["P"], dd["g"], *_d)
    qn = npd.ComputeDarcyFlux(d["rhofv"], *d["kP"], d["P"], dd["g"], *_d)
    for a, b in zip(q, qn):
        assert np.array_equal(a, b)
    orc.ResidualFluidPressure(F, d["phc"], d["phirho"], d["phirhof0"], PC, *q, 1.0, 1 / dd["dt"], *_d)
    assert np.array_equal(F, npd.ResidualFluidPressure(d["phc"], d["phirho"], d["phirhof0"], PC, *qn, 1.0,
                                                       1 / dd["dt"], *_d))


def _heat_residual(npd, d, T):
    h = d["heat"]
    _d = (1 / h["dx"], 1 / h["dy"], 1 / h["dz"])
    Tb = npd.SetTemperatureBCs_serial(T, d["kT"][1], d["wN"], d["phc"], h["qyS"], _d[1], h["TN"])
    q = npd.ComputeFlux(*d["kT"], Tb, *_d)
    return npd.ResidualTemperatureLinearised(Tb, d["T0"], d["rhoC"], d["phc"], 1.0, *q, h["Qt"], h["transient"],
                                             1 / h["dt"], *_d)


@pytest.mark.parametrize("immersed", [False, True], ids=["flat", "immersed"])
def test_v18_heat_solve_matches_dense_solve(orc, npd, immersed):
    """The converged state of the oracle's v18 heat loop equals the dense solve
    of the affine residual assembled from the NumPy restatement (serial ghost
    fill, fluxes, linearised residual) on the cells that carry an equation."""
    grid = (6, 8, 4)
    d = _v18(grid, immersed=immersed, dt_s=1e14)
    if immersed:
        d["wN"][:] = np.where(d["wN"] < 1.0, 0.9, 1.0)    # a stable extrapolation weight
    free = np.abs(d["phc"][1:-1, 1:-1, 1:-1]) != 1.0
    idx = np.flatnonzero(free.ravel(order="F"))
    base = d["T"].copy(order="F")

    def apply(x):
        T = base.copy(order="F")
        inner = T[1:-1, 1:-1, 1:-1].ravel(order="F")
        inner[idx] = x
        T[1:-1, 1:-1, 1:-1] = inner.reshape(grid, order="F")
        return _heat_residual(npd, d, T).ravel(order="F")[idx]

    f0 = apply(np.zeros(len(idx)))
    M = np.column_stack([apply(np.eye(len(idx))[:, c]) - f0 for c in range(len(idx))])
    x = np.linalg.solve(M, -f0)
    T = base.copy(order="F")
    rep, _ = orc.v18_thermal_solve(T, d["T0"], d["rhoC"], *d["kT"], d["phc"], d["wN"], **d["heat"],
                                   p=orc.params(nitmax=100000, nitout=100, ncheck=10, norm=1, tol=1e-13))
    assert rep["status"] == 0
    got = T[1:-1, 1:-1, 1:-1].ravel(order="F")[idx]
    assert np.max(np.abs(got - x)) <= 1e-10 * np.max(np.abs(x))


@pytest.mark.parametrize("immersed", [False, True], ids=["flat", "immersed"])
def test_v18_darcy_solve_zeroes_the_residual(orc, npd, immersed):
    """The Darcy loop is nonlinear (the fluid density depends on P): its
    converged state makes the NumPy restatement's residual (serial ghost fill,
    EOS, Darcy flux with gravity) vanish relative to the initial one."""
    grid = (9, 8, 5)
    d = _v18(grid, immersed=immersed, dt_s=1e14)
    d["wN"][:] = np.where(d["wN"] < 1.0, 0.9, 1.0)
    dd = d["darcy"]
    _d = (1 / dd["dx"], 1 / dd["dy"], 1 / dd["dz"])

    def residual(P):
        Pb = npd.SetPressureBCs_serial(P, d["wN"], d["phc"], dd["PN"])
        prf = npd.PhiRhof(np.zeros(grid, order="F"), d["T"], Pb, d["phc"], dd["phi_i"], dd["scale_rho"],
                          dd["scale_T"], dd["scale_sigma"])
        q = npd.ComputeDarcyFlux(d["rhofv"], *d["kP"], Pb, dd["g"], *_d)
        return npd.ResidualFluidPressure(d["phc"], prf, d["phirhof0"], 1.0, *q, dd["transient"], 1 / dd["dt"], *_d)

    r0 = np.max(np.abs(residual(d["P"])))
    P, phirho = d["P"].copy(order="F"), d["phirho"].copy(order="F")
    rep, _ = orc.v18_darcy_solve(P, d["T"], phirho, d["phirhof0"], d["rhofv"], *d["kP"], d["phc"], d["wN"],
                                 **dd, p=orc.params(nitmax=100000, nitout=100, ncheck=10, norm=1, tol=1e-13))
    assert rep["status"] == 0 and rep["iters"] > 50
    assert np.max(np.abs(residual(P))) <= 1e-10 * r0


# ------------------------------------------------------------ output fields
def _eq_nan(a, b):
    return np.array_equal(a, b, equal_nan=True)


def test_output_fields_match_numpy(orc, npd):
    grid = (9, 7, 6)
    d = dim_state(*grid, np.random.default_rng(31), air_from=5)
    sc, sh = d["sc"], d["sh"]
    rng = np.random.default_rng(32)
    z = lambda s: np.zeros(sh[s], order="F")
    V = [np.asfortranarray(rng.uniform(-1, 1, sh[s])) for s in ("Fx", "Fy", "Fz")]
    X = z("c")
    orc.Phase(X, d["phc"])
    assert _eq_nan(X, npd.Phase(d["phc"]))
    orc.Pressure(X, d["P"], d["phc"], sc.sigma)
    assert _eq_nan(X, npd.Scaled(d["P"], d["phc"], sc.sigma))
    orc.Temperature(X, d["T"], d["phc"], sc.T)
    assert _eq_nan(X, npd.Scaled(d["T"], d["phc"], sc.T))
    XYZ0 = [np.asfortranarray(rng.uniform(-1, 1, sh["c"])) for _ in range(3)]
    XYZ = [a.copy(order="F") for a in XYZ0]
    orc.Velocity(*XYZ, *V, d["phc"], sc.V)
    for a, b in zip(XYZ, npd.Velocity(*XYZ0, *V, d["phc"], sc.V)):
        assert _eq_nan(a, b)
    orc.EffectiveThermalConductivity(X, d["T"], 0.01, d["phc"], sc.T)
    assert _eq_nan(X, npd.EffectiveThermalConductivity(d["T"], 0.01, d["phc"], sc.T))
    C0 = np.asfortranarray(rng.uniform(-1, 1, sh["c"]))
    C = C0.copy(order="F")
    orc.EffectiveHeatCapacity(C, d["T"], d["P"], d["phc"], 0.01, sc.T, sc.sigma, sc.C, sc.rho)
    assert _eq_nan(C, npd.EffectiveHeatCapacity(C0, d["T"], d["phc"], sc.T))
    orc.Permeability(X, 50.0, -30e3 / sc.L, 5.0, d["phc"], 3000.0 / sc.L, sc.L)
    np.testing.assert_allclose(X, npd.Permeability(50.0, -30e3 / sc.L, 5.0, d["phc"], 3000.0 / sc.L, sc.L, sh["c"]),
                               rtol=1e-15)
    orc.FluidDensity(X, d["T"], d["P"], d["phc"], sc.T, sc.sigma)
    assert _eq_nan(X, npd.FluidDensity(d["T"], d["P"], d["phc"], sc.T, sc.sigma))
    orc.Viscosity(X, d["T"], d["phc"], sc.T)
    np.testing.assert_allclose(X, npd.Viscosity(d["T"], d["phc"], sc.T), rtol=1e-15)
    orc.Peclet(X, d["T"], d["P"], *V, d["phc"], 0.01, sc.T, sc.sigma, sc.V, sc.L)
    np.testing.assert_allclose(X, npd.Peclet(d["T"], d["P"], *V, d["phc"], 0.01, sc.T, sc.sigma, sc.V, sc.L),
                               rtol=1e-15)


def test_output_fields_closed_forms(orc):
    """Uniform flow and temperature: the Peclet number is L rho_f C_f |V| / k_eff
    from the closures; the velocity field is the face value; air is NaN."""
    from hydrothermal3d_amd.dimensional import v18_scaling
    sc = v18_scaling()
    grid = (5, 4, 3)
    sh = shapes(*grid)
    T = np.full(sh["E"], 500.0 / sc.T, order="F")
    P = np.full(sh["E"], 2e7 / sc.sigma, order="F")
    ph = np.zeros(sh["E"], order="F")
    ph[:, -2:, :] = 1.0
    v = (3e-9 / sc.V, -4e-9 / sc.V, 12e-9 / sc.V)
    V = [np.full(sh[s], vv, order="F") for s, vv in zip(("Fx", "Fy", "Fz"), v)]
    X = np.zeros(sh["c"], order="F")
    orc.Peclet(X, T, P, *V, ph, 0.01, sc.T, sc.sigma, sc.V, sc.L)
    rho = orc.eos("rhof_C", 500.0 - 273.15, 2e7)
    want = sc.L * rho * (1000 + 7.5 * 500.0) * 13e-9 / (0.99 * (3.1138 - 0.0023 * 500) + 0.01 * (-0.069 + 0.0012 * 500))
    rock = ph[1:-1, 1:-1, 1:-1] != 1.0
    np.testing.assert_allclose(X[rock], want, rtol=1e-12)
    assert np.all(np.isnan(X[~rock]))
    Y = [np.zeros(sh["c"], order="F") for _ in range(3)]
    orc.Velocity(*Y, *V, ph, sc.V)
    for a, vv in zip(Y, (3e-9, -4e-9, 12e-9)):
        np.testing.assert_allclose(a[:, :2, :], vv, rtol=1e-14)
