"""Checks of the CPU oracle that the mathematics fixes (runs without a GPU).

Each test feeds an oracle kernel an input whose exact output follows from the
discretisation itself -- linear and quadratic fields on which the staggered
averages and differences are exact, hydrostatic states, geometric series of
the damped updates, the Gershgorin row sums of the assembled 7-point matrix,
the stability limit of the damped second-order iteration, the covariance of
the spectral random field -- chosen so that a plausible slip (a wrong
averaging plane, index, sign, scale factor or transposed operand) fails.
Coefficients are dyadic where the expected values are compared exactly.
"""
import math

import numpy as np
import pytest

from helpers import shapes

# integer-valued slopes, different per axis: a permuted axis or shifted plane changes the result
SX, SY, SZ, C0 = 1.0, 16.0, 256.0, 3.0


def F(shape):
    return np.zeros(shape, order="F")


def fa(a):
    return np.asfortranarray(a)


def _lin(shape, off):
    """L(p) = C0 + SX p_x + SY p_y + SZ p_z at grid positions p = index + off
    (index units: vertices at integers, cell centres at half-integers)."""
    i, j, k = np.meshgrid(*(np.arange(n) + o for n, o in zip(shape, off)), indexing="ij")
    return fa(C0 + SX * i + SY * j + SZ * k)


# ------------------------------------------------------------------ face averages
@pytest.mark.parametrize("kernel", ["InitThermal", "InitDarcy", "InitConductivity"])
def test_face_averages_of_a_linear_vertex_field(orc, kernel):
    """kernels_HT3D.jl:82-107, kernels_HT3D_dimensional.jl:133-138: @av_yza,
    @av_xza and @av_xya average the four vertices around a face, which for a
    linear field is its value at the face centre: kx at (i, j+1/2, k+1/2), ky
    at (i+1/2, j, k+1/2), kz at (i+1/2, j+1/2, k)."""
    nx, ny, nz = 5, 4, 3
    sh = shapes(nx, ny, nz)
    kv = _lin(sh["V"], (0, 0, 0))
    kx, ky, kz = F(sh["Fx"]), F(sh["Fy"]), F(sh["Fz"])
    T = _lin(sh["E"], (-0.5, -0.5, -0.5))
    if kernel == "InitThermal":
        Rt = F(sh["c"])
        orc.InitThermal(Rt, kx, ky, kz, T, kv, 4.0)
        assert np.array_equal(Rt, -4.0 * T[1:-1, 1:-1, 1:-1])
    elif kernel == "InitDarcy":
        Ty = F(sh["Fy"])
        orc.InitDarcy(Ty, kx, ky, kz, T, kv, 1.0)
        # @av_yi(Tc_ex): the cell-centre field at the y face (i+1/2, j, k+1/2)
        assert np.array_equal(Ty, _lin(sh["Fy"], (0.5, 0.0, 0.5)))
    else:
        orc.InitConductivity(kx, ky, kz, kv)
    assert np.array_equal(kx, _lin(sh["Fx"], (0.0, 0.5, 0.5)))
    assert np.array_equal(ky, _lin(sh["Fy"], (0.5, 0.0, 0.5)))
    assert np.array_equal(kz, _lin(sh["Fz"], (0.5, 0.5, 0.0)))


# ------------------------------------------------------------------ UpdateT!
def test_updatet_conductive_profile_is_a_fixed_point(orc):
    """kernels_HT3D.jl:119-126 with R = -T/dt (InitThermal!): on T = 1 - y with
    Dirichlet ghosts the flux is uniform, so F = R + T/dt + div q = 0 and T
    does not move."""
    nx, ny, nz = 4, 8, 4
    dy = 1.0 / ny
    sh = shapes(nx, ny, nz)
    yce = (np.arange(ny + 2) - 0.5) * dy
    T = fa(np.broadcast_to((1.0 - yce)[None, :, None], sh["E"]).copy())
    _dt = 64.0
    R = fa(-_dt * T[1:-1, 1:-1, 1:-1])
    k = [np.ones(sh[s], order="F") for s in ("Fx", "Fy", "Fz")]
    q = [F(sh[s]) for s in ("Fx", "Fy", "Fz")]
    orc.ComputeFlux(*q, *k, T, 4.0, 1 / dy, 4.0)
    Fr, T1 = F(sh["c"]), T.copy(order="F")
    orc.UpdateT(Fr, T1, R, *q, _dt, 0.01, 4.0, 1 / dy, 4.0)
    assert np.max(np.abs(Fr)) < 1e-11
    assert np.max(np.abs(T1 - T)) < 1e-13


def test_updatet_point_source_gives_the_7_point_stencil(orc):
    """A unit value in one interior cell (zero elsewhere, R = 0): F at that cell
    is 1/dt + 2 (kx/dx^2 + ky/dy^2 + kz/dz^2), at its six neighbours -k/h^2 of
    the shared face, zero elsewhere; T_inn -= dtau F, ghosts untouched."""
    nx, ny, nz = 5, 6, 7
    sh = shapes(nx, ny, nz)
    _dx, _dy, _dz, _dt, dtau = 4.0, 8.0, 2.0, 16.0, 1.0 / 64
    kxv, kyv, kzv = 2.0, 0.5, 4.0
    T = F(sh["E"])
    c = (2, 3, 4)                      # interior cell, 0-based cell index
    T[c[0] + 1, c[1] + 1, c[2] + 1] = 1.0
    kx, ky, kz = (np.full(sh[s], v, order="F") for s, v in (("Fx", kxv), ("Fy", kyv), ("Fz", kzv)))
    q = [F(sh[s]) for s in ("Fx", "Fy", "Fz")]
    orc.ComputeFlux(*q, kx, ky, kz, T, _dx, _dy, _dz)
    Fr, T1 = F(sh["c"]), T.copy(order="F")
    orc.UpdateT(Fr, T1, F(sh["c"]), *q, _dt, dtau, _dx, _dy, _dz)
    want = F(sh["c"])
    want[c] = _dt + 2 * (kxv * _dx**2 + kyv * _dy**2 + kzv * _dz**2)
    for ax, (kk, ih) in enumerate(((kxv, _dx), (kyv, _dy), (kzv, _dz))):
        for s in (-1, 1):
            n = list(c)
            n[ax] += s
            want[tuple(n)] = -kk * ih**2
    assert np.array_equal(Fr, want)
    Tw = T.copy(order="F")
    Tw[1:-1, 1:-1, 1:-1] -= dtau * want
    assert np.array_equal(T1, Tw)


# ------------------------------------------------------------------ damped updates
def test_updatep_is_a_geometric_series(orc):
    """kernels_HT3D.jl:136-142 (= DampedUpdate!, kernels_HT3D_dimensional.jl:316-320):
    with a constant F and F0 = 0, n updates give F0 = F (1 - d^n) / (1 - d)
    and move P by -dtau F sum_{m=1..n} (1 - d^m) / (1 - d); the ghost layer
    is not written."""
    nx, ny, nz = 4, 3, 5
    sh = shapes(nx, ny, nz)
    d, dtau = 0.75, 0.125
    rng = np.random.default_rng(3)
    Fr = fa(rng.integers(-8, 9, sh["c"]).astype(float))
    P = fa(rng.integers(-8, 9, sh["E"]).astype(float))
    P0 = P.copy()
    F0 = F(sh["c"])
    n = 5
    for _ in range(n):
        orc.UpdateP(F0, P, Fr, d, dtau)
    geo = [(1 - d**m) / (1 - d) for m in range(1, n + 1)]
    assert np.allclose(F0, Fr * geo[-1], rtol=1e-15, atol=0)
    shift = P0.copy()
    shift[1:-1, 1:-1, 1:-1] -= dtau * Fr * sum(geo)
    assert np.allclose(P, shift, rtol=1e-14, atol=1e-14)
    g = np.ones(sh["E"], dtype=bool)
    g[1:-1, 1:-1, 1:-1] = False
    assert np.array_equal(P[g], P0[g])
    # two steps, the first checked exactly: F0 = F, then F0 = F (1 + d)
    F0 = F(sh["c"])
    P = P0.copy(order="F")
    orc.UpdateP(F0, P, Fr, d, dtau)
    assert np.array_equal(F0, Fr)
    orc.UpdateP(F0, P, Fr, d, dtau)
    assert np.array_equal(F0, Fr * (1 + d))


def test_dyrelupdate_is_a_geometric_series(orc):
    """kernels_HT3D_dimensional.jl:322-326: with a constant F and F0 = 0,
    F0_n = h2 F (1 - h1^n) / (1 - h1), and X moves by the partial sums."""
    sh = shapes(3, 4, 5)
    h1, h2, dtau = 0.5, 0.25, 2.0
    rng = np.random.default_rng(4)
    Fr = fa(rng.integers(-8, 9, sh["c"]).astype(float))
    X = F(sh["E"])
    F0 = F(sh["c"])
    n = 6
    for _ in range(n):
        orc.DYRELUpdate(F0, X, Fr, h1, h2, dtau)
    geo = [h2 * (1 - h1**m) / (1 - h1) for m in range(1, n + 1)]
    assert np.array_equal(F0, Fr * geo[-1])          # dyadic: exact
    assert np.array_equal(X[1:-1, 1:-1, 1:-1], -dtau * Fr * sum(geo))


def test_dyrel_stability_limit(orc):
    """The DYREL scalars (HydroThermal3D_v18.jl:433-439) with the update of
    kernels_HT3D_dimensional.jl:322-326 form, on an eigenmode of the
    preconditioned operator with eigenvalue mu, the 2x2 iteration
    [F0; a] <- [[h1, h2 mu], [-dtau h1, 1 - dtau h2 mu]] [F0; a], whose
    determinant is h1 and whose trace is 1 + h1 - dtau h2 mu.  By the Jury
    criterion it is stable for every damping c > 0 iff dtau^2 mu < 4, i.e.
    with dtau = 2 CFL / sqrt(lambda_max): iff CFL < sqrt(lambda_max / mu).
    On an nx x 1 x 1 Neumann grid (uniform k, no source) the highest x mode
    cos(pi m (i + 1/2) / nx), m = nx - 1, is the top of the spectrum, with
    mu = (2 - 2 cos(pi m / nx)) / dx^2 / D, D = 2/dx^2 + 2/dy^2 + 2/dz^2 the
    Gershgorin diagonal (lambda_max = 2).  Just below that CFL the mode decays,
    just above it grows.  A dropped square root in dtau, a missing factor 2 in
    h2 or a wrong D moves the limit and fails one side."""
    nx, dx, dy, dz = 8, 0.1, 1.0, 1.0
    m = nx - 1
    D = 2 / dx**2 + 2 / dy**2 + 2 / dz**2
    mu = (2 - 2 * math.cos(math.pi * m / nx)) / dx**2 / D
    cfl_star = math.sqrt(2.0 / mu)
    sh = shapes(nx, 1, 1)
    k = [np.ones(sh[s], order="F") for s in ("Fx", "Fy", "Fz")]
    i = np.arange(nx)
    mode = np.cos(np.pi * m * (i + 0.5) / nx)[:, None, None]
    out = {}
    for label, cfl in (("below", 0.99 * cfl_star), ("above", 1.01 * cfl_star)):
        X = F(sh["E"])
        X[1:-1, 1:-1, 1:-1] = mode
        rep, _ = orc.dyrel_solve(X, F(sh["c"]), *k, dx=dx, dy=dy, dz=dz, dt=1.0, transient=0, rhoC=0.0,
                                 p=orc.params(nitmax=3000, nitout=10**6, ncheck=100, norm=1, tol=0.0, CFL=cfl))
        assert rep["lmax"] == 2.0
        assert abs(rep["dtau"] ** 2 * mu - 4 * (cfl / cfl_star) ** 2) < 1e-12
        out[label] = (rep, np.max(np.abs(X[1:-1, 1:-1, 1:-1])))
    rb, ab = out["below"]
    ra, aa = out["above"]
    assert rb["status"] == 10 and ab < 1e-3 and rb["nF_rel"] < 1e-3       # ran to nitmax, decayed
    assert ra["status"] == 12 or aa > 10.0                                  # diverged


# ------------------------------------------------------------------ Gershgorin
def _assembled_rows(orc, k, h, rho_tr_dt, grid):
    """The 7-point matrix A (+ diag rho*transient/dt) assembled column by column
    from the oracle's own flux and divergence kernels on a Neumann grid."""
    sh = shapes(*grid)
    N = int(np.prod(grid))
    cols = []
    for c in range(N):
        E = F(sh["E"])
        E[1:-1, 1:-1, 1:-1] = np.eye(N)[:, c].reshape(grid, order="F")
        orc.SetPressureBCs(E, 0.0, 0.0)
        q = [F(sh[s]) for s in ("Fx", "Fy", "Fz")]
        orc.ComputeFlux(*q, *k, E, 1 / h[0], 1 / h[1], 1 / h[2])
        r = F(sh["c"])
        orc.ResidualDiffusion(r, F(sh["c"]), *q, 1 / h[0], 1 / h[1], 1 / h[2])
        cols.append(r.ravel(order="F") + rho_tr_dt * np.eye(N)[:, c])
    return np.column_stack(cols)


@pytest.mark.parametrize("rho", [0.5, -0.25, 0.0])
def test_gershgorin_is_the_row_sum_of_the_assembled_matrix(orc, rho):
    """GershgorinPoisson! (kernels_HT3D_dimensional.jl:207-221) and
    GershgorinHydro! (:335-347) against the Gershgorin disc of the assembled
    7-point matrix: for a cell whose six faces are all interior, the sum of
    |row| is GershgorinHydro's G, the diagonal is GershgorinPoisson's D and
    their ratio its G (random k, three different spacings, a capacity term of
    either sign)."""
    grid = (5, 4, 6)
    sh = shapes(*grid)
    rng = np.random.default_rng(11)
    k = [fa(10.0 ** rng.uniform(-1, 1, sh[s])) for s in ("Fx", "Fy", "Fz")]
    h, dt, tr = (0.2, 0.25, 1 / 6), 0.05, 1.0
    A = _assembled_rows(orc, k, h, rho * tr / dt, grid)
    rhoC = np.full(sh["c"], rho, order="F")
    G, D, GH = F(sh["c"]), F(sh["c"]), F(sh["c"])
    orc.GershgorinPoisson(G, D, rhoC, *k, tr, dt, *h)
    orc.GershgorinHydro(GH, rhoC, *k, tr, dt, *h)
    idx = np.arange(A.shape[0]).reshape(grid, order="F")[1:-1, 1:-1, 1:-1].ravel(order="F")
    rows = np.abs(A[idx]).sum(axis=1)
    diag = A[idx, idx]
    Gi = lambda a: a[1:-1, 1:-1, 1:-1].ravel(order="F")
    assert np.allclose(Gi(GH), rows, rtol=1e-13)
    assert np.allclose(Gi(D), diag, rtol=1e-13)
    assert np.allclose(Gi(G), rows / diag, rtol=1e-13)


def test_gershgorin_hydro_uniform(orc):
    """GershgorinHydro! with uniform k and h: G = 6k/h^2 + |6k/h^2 + s/dt|."""
    sh = shapes(4, 5, 3)
    h, dt, kv = 0.125, 0.25, 2.0
    k = [np.full(sh[s], kv, order="F") for s in ("Fx", "Fy", "Fz")]
    for s in (3.0, -1000.0):
        G = F(sh["c"])
        orc.GershgorinHydro(G, np.full(sh["c"], s, order="F"), *k, 1.0, dt, h, h, h)
        assert np.all(G == 6 * kv / h**2 + abs(6 * kv / h**2 + s / dt))


# ------------------------------------------------------------------ Darcy flux and velocity
def test_darcy_flux_hydrostatic_state(orc):
    """ComputeDarcyFlux! (kernels_HT3D_dimensional.jl:266-272): with a fluid
    density linear in x and z and a column-wise hydrostatic pressure
    P = g rho(x_c, z_c) y_c, the vertical flux -ky (dP/dy - g <rho>_y) vanishes
    and the lateral fluxes are -k g y_c d(rho)/dx and -k g y_c d(rho)/dz."""
    nx, ny, nz = 4, 5, 3
    sh = shapes(nx, ny, nz)
    dx, dy, dz, g = 0.5, 0.25, 0.125, 8.0
    a, bx, bz = 2.0, 0.5, 0.25
    iv, jv, kv = np.meshgrid(*(np.arange(n) for n in sh["V"]), indexing="ij")
    rhofv = fa(a + bx * iv * dx + bz * kv * dz + 0 * jv)
    ie, je, ke = np.meshgrid(*(np.arange(n) - 0.5 for n in sh["E"]), indexing="ij")
    P = fa(g * (a + bx * ie * dx + bz * ke * dz) * (je * dy))
    kx, ky, kz = (np.full(sh[s], v, order="F") for s, v in (("Fx", 2.0), ("Fy", 4.0), ("Fz", 0.5)))
    q = [F(sh[s]) for s in ("Fx", "Fy", "Fz")]
    orc.ComputeDarcyFlux(*q, rhofv, kx, ky, kz, P, g, 1 / dx, 1 / dy, 1 / dz)
    assert np.max(np.abs(q[1])) < 1e-13
    yc = (np.arange(ny) + 0.5) * dy
    assert np.allclose(q[0], -2.0 * g * bx * yc[None, :, None] * np.ones(sh["Fx"]), rtol=1e-13)
    assert np.allclose(q[2], -0.5 * g * bz * yc[None, :, None] * np.ones(sh["Fz"]), rtol=1e-13)


def test_init_vel_buoyancy_balance(orc):
    """Init_vel! (kernels_HT3D.jl:73-80): a pressure that balances the buoyancy,
    P = Ra tau y with Ty = tau, gives Vy = 0; a pressure linear in x gives
    Vx = -kx dP/dx."""
    nx, ny, nz = 3, 4, 5
    sh = shapes(nx, ny, nz)
    dx, dy, dz, Ra, tau = 0.25, 0.125, 0.5, 64.0, 0.75
    ie, je, ke = np.meshgrid(*(np.arange(n) - 0.5 for n in sh["E"]), indexing="ij")
    P = fa(Ra * tau * je * dy + 3.0 * ie * dx)
    k = [np.full(sh[s], 2.0, order="F") for s in ("Fx", "Fy", "Fz")]
    V = [F(sh[s]) for s in ("Fx", "Fy", "Fz")]
    orc.Init_vel(*V, *k, P, np.full(sh["Fy"], tau, order="F"), Ra, 1 / dx, 1 / dy, 1 / dz)
    assert np.array_equal(V[1], F(sh["Fy"]))
    assert np.array_equal(V[0], np.full(sh["Fx"], -6.0))
    assert np.array_equal(V[2], F(sh["Fz"]))


def test_compute_rp_linear_buoyancy(orc):
    """Compute_Rp! (kernels_HT3D.jl:92-97, @dmul_ya): uniform ky and Ty linear in
    y give Rp = Ra ky dTy/dy, uniform over the cells."""
    nx, ny, nz = 3, 6, 4
    sh = shapes(nx, ny, nz)
    dy, Ra = 0.125, 16.0
    Ty = fa(np.broadcast_to((2.0 - 0.5 * np.arange(ny + 1) * dy)[None, :, None], sh["Fy"]).copy())
    Rp = F(sh["c"])
    orc.Compute_Rp(Rp, np.full(sh["Fy"], 4.0, order="F"), Ty, Ra, 1 / dy)
    assert np.array_equal(Rp, np.full(sh["c"], Ra * 4.0 * -0.5))


# ------------------------------------------------------------------ dimensional residuals
def _phases(sh, rng):
    ph = fa(rng.choice([0.0, 2.0, 3.0, 1.0, -1.0], size=sh["E"]))
    return ph


def test_residual_temperature_linearised_manufactured(orc):
    """ResidualTemperatureLinearised! (kernels_HT3D_dimensional.jl:162-181) on
    T = T_lin + x^2 with uniform k: the 7-point divergence of -k grad T is
    exactly -2k (per unit x-spacing squared), so F = (rhoC (T - T0)/dt - 2k - Q)/PC
    with Q = Qt only in phase-3 cells, and F = 0 in air and immersed cells."""
    nx, ny, nz = 6, 5, 4
    sh = shapes(nx, ny, nz)
    rng = np.random.default_rng(5)
    dx, dy, dz, kv, _dt, Qt = 0.5, 0.25, 0.125, 2.0, 4.0, 3.0
    ie, je, ke = np.meshgrid(*(np.arange(n) - 0.5 for n in sh["E"]), indexing="ij")
    T = fa(1.0 + 0.5 * je * dy + 0.25 * ke * dz + (ie * dx) ** 2)
    T0 = fa(T[1:-1, 1:-1, 1:-1] - rng.integers(-4, 5, sh["c"]) * 0.125)
    rhoC = fa(rng.integers(1, 5, sh["c"]).astype(float))
    PC = fa(2.0 ** rng.integers(-2, 3, sh["c"]))
    ph = _phases(sh, rng)
    k = [np.full(sh[s], kv, order="F") for s in ("Fx", "Fy", "Fz")]
    q = [F(sh[s]) for s in ("Fx", "Fy", "Fz")]
    orc.ComputeFlux(*q, *k, T, 1 / dx, 1 / dy, 1 / dz)
    for transient in (1.0, 0.0):
        Fr = F(sh["c"])
        orc.ResidualTemperatureLinearised(Fr, T, T0, rhoC, ph, PC, *q, Qt, transient, _dt, 1 / dx, 1 / dy, 1 / dz)
        phi = ph[1:-1, 1:-1, 1:-1]
        Q = np.where(phi == 3.0, Qt, 0.0)
        want = (transient * rhoC * _dt * (T[1:-1, 1:-1, 1:-1] - T0) - 2 * kv - transient * Q) / PC
        want[np.abs(phi) == 1.0] = 0.0
        assert np.allclose(Fr, want, rtol=1e-13, atol=1e-12)
        assert np.all(Fr[np.abs(phi) == 1.0] == 0.0)


def test_residual_fluid_pressure_manufactured(orc):
    """ResidualFluidPressure! (kernels_HT3D_dimensional.jl:274-289) with the
    Darcy flux of a hydrostatic pressure plus x^2 (uniform k and density):
    the flux divergence is exactly -2k, so F = ((phirho - phirho0)/dt - 2k)/PC,
    and F = 0 in air and immersed cells."""
    nx, ny, nz = 5, 6, 4
    sh = shapes(nx, ny, nz)
    rng = np.random.default_rng(6)
    dx, dy, dz, g, rho, kv, _dt = 0.5, 0.25, 0.125, 8.0, 2.0, 4.0, 16.0
    ie, je, ke = np.meshgrid(*(np.arange(n) - 0.5 for n in sh["E"]), indexing="ij")
    P = fa(g * rho * je * dy + (ie * dx) ** 2)
    k = [np.full(sh[s], kv, order="F") for s in ("Fx", "Fy", "Fz")]
    q = [F(sh[s]) for s in ("Fx", "Fy", "Fz")]
    orc.ComputeDarcyFlux(*q, np.full(sh["V"], rho, order="F"), *k, P, g, 1 / dx, 1 / dy, 1 / dz)
    pr = fa(rng.integers(0, 9, sh["c"]) * 0.25)
    pr0 = fa(rng.integers(0, 9, sh["c"]) * 0.25)
    PC = fa(2.0 ** rng.integers(-2, 3, sh["c"]))
    ph = _phases(sh, rng)
    for transient in (1.0, 0.0):
        Fr = F(sh["c"])
        orc.ResidualFluidPressure(Fr, ph, pr, pr0, PC, *q, transient, _dt, 1 / dx, 1 / dy, 1 / dz)
        phi = ph[1:-1, 1:-1, 1:-1]
        want = (transient * _dt * (pr - pr0) - 2 * kv) / PC
        want[np.abs(phi) == 1.0] = 0.0
        assert np.allclose(Fr, want, rtol=1e-13, atol=1e-12)
        assert np.all(Fr[np.abs(phi) == 1.0] == 0.0)


# ------------------------------------------------------------------ SwapDYREL!
def test_swapdyrel_copies_both_outputs(orc):
    """SwapDYREL! (kernels_HT3D_dimensional.jl:122-126): Xc = inn(X0c_ex) and Yc = Y0c."""
    sh = shapes(4, 3, 5)
    rng = np.random.default_rng(1)
    X0 = fa(rng.standard_normal(sh["E"]))
    Y0 = fa(rng.standard_normal(sh["c"]))
    Xc, Yc = fa(rng.standard_normal(sh["c"])), fa(rng.standard_normal(sh["c"]))
    orc.SwapDYREL(Xc, Yc, X0, Y0)
    assert np.array_equal(Xc, X0[1:-1, 1:-1, 1:-1]) and np.array_equal(Yc, Y0)


# ------------------------------------------------------------------ synthetic permeability
def test_lognormal_field_statistics(orc):
    """The seeded permeability generator: log10 k = sigma G with G a Gaussian
    field of unit variance and exponential covariance exp(-r/l).  Over 12 seeds
    on 41^3 vertices of the unit cube (l = 0.1, sigma = 1) the sample mean of
    log10 k is 0 and its variance 1 within sampling error, and the lagged
    products match exp(-r/l) at r = l/2, l and 2l; a Gaussian covariance, a
    doubled l or sigma applied to ln k would each fall outside."""
    from hydrothermal3d_amd import inputs
    n, ell = 40, 0.1
    stats = []
    for s in range(12):
        w = inputs.Workload("stats", n, n, n, sigma=1.0, seed=700 + s, corr_len=ell)
        kfv = np.ones((n + 1,) * 3, order="F")
        orc.lognormal_field(kfv, *inputs.field_args(w))
        L = np.log10(kfv)
        row = [L.mean(), np.mean(L * L)]
        for lag in (2, 4, 8):
            row.append(np.mean([np.mean(L[lag:] * L[:-lag]), np.mean(L[:, lag:] * L[:, :-lag]),
                                np.mean(L[:, :, lag:] * L[:, :, :-lag])]))
        stats.append(row)
    m = np.mean(stats, axis=0)
    assert abs(m[0]) < 0.1
    assert abs(m[1] - 1.0) < 0.08
    for lag, got in zip((2, 4, 8), m[2:]):
        assert abs(got - math.exp(-lag / n / ell)) < 0.06, (lag, got)
