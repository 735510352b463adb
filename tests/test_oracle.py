"""Pinning the CPU oracle (runs without a GPU).

The reference (Julia + ParallelStencil) cannot run here and ships no tests or
fixtures, so the oracle is pinned by
  * a second, independent restatement (oracle/ht3d_numpy.py, array slicing
    instead of point loops) that must agree bit for bit kernel by kernel;
  * closed forms and textbook special cases a dropped term, wrong sign, wrong
    index or transposed operand would break;
  * brute force: dense solves of the assembled 7-point systems on tiny grids;
  * invariants (compatibility, symmetry, mean preservation, constant fields);
  * the reference's own known-answer check, tools/TestAdvection.jl:84-117.
"""
import math

import numpy as np
import pytest

from helpers import shapes, rand, positive
from oracle import ht3d_numpy as npo

GRIDS = [(3, 3, 3), (7, 5, 6), (13, 9, 11)]


def F(shape):
    return np.zeros(shape, order="F")


def fa(a):
    return np.asfortranarray(a)


# ---------------------------------------------------------------- independent restatements agree
@pytest.mark.parametrize("grid", GRIDS, ids=lambda g: "x".join(map(str, g)))
def test_coefficients_numpy_vs_c(orc, grid):
    rng = np.random.default_rng(1)
    sh = shapes(*grid)
    T, kv = rand(sh["E"], rng), positive(sh["V"], rng)
    Rt, kx, ky, kz = F(sh["c"]), F(sh["Fx"]), F(sh["Fy"]), F(sh["Fz"])
    orc.InitThermal(Rt, kx, ky, kz, T, kv, 37.0)
    for a, b in zip((Rt, kx, ky, kz), npo.InitThermal(T, kv, 37.0)):
        assert np.array_equal(a, b)
    Ty = F(sh["Fy"])
    orc.InitDarcy(Ty, kx, ky, kz, T, kv, 1.0)
    for a, b in zip((Ty, kx, ky, kz), npo.InitDarcy(T, kv)):
        assert np.array_equal(a, b)
    orc.KillNorthSouthFluxes(Ty)
    Tyn = npo.KillNorthSouthFluxes(npo.InitDarcy(T, kv)[0])
    assert np.array_equal(Ty, Tyn)
    Rp = F(sh["c"])
    orc.Compute_Rp(Rp, ky, Ty, 317.0, 9.5)
    assert np.array_equal(Rp, npo.Compute_Rp(ky, Tyn, 317.0, 9.5))


@pytest.mark.parametrize("grid", GRIDS, ids=lambda g: "x".join(map(str, g)))
def test_viscosity_numpy_vs_c(orc, grid):
    """Face mobility k exp(a T_face); libm exp on both sides of the comparison."""
    rng = np.random.default_rng(12)
    sh = shapes(*grid)
    T = rand(sh["E"], rng, 0.0, 1.0)
    k = [positive(sh[s], rng) for s in ("Fx", "Fy", "Fz")]
    kc = [a.copy(order="F") for a in k]
    orc.ApplyViscosity(*kc, T, math.log(10.0))
    for a, b in zip(kc, npo.ApplyViscosity(*k, T, math.log(10.0))):
        assert np.allclose(a, b, rtol=6e-16, atol=0)   # numpy and libm exp differ by an ulp
    # uniform T: a pure factor 10^T
    T[...] = 0.25
    kc = [a.copy(order="F") for a in k]
    orc.ApplyViscosity(*kc, T, math.log(10.0))
    for a, b in zip(kc, k):
        assert np.allclose(a, b * 10 ** 0.25, rtol=1e-15)


@pytest.mark.parametrize("grid", GRIDS, ids=lambda g: "x".join(map(str, g)))
def test_bcs_numpy_vs_c(orc, grid):
    rng = np.random.default_rng(2)
    sh = shapes(*grid)
    A = rand(sh["E"], rng)
    B = A.copy(order="F")
    orc.SetTemperatureBCs(B, 1.0, 0.25)
    assert np.array_equal(B, npo.temperature_bc(A, 1.0, 0.25))
    B = A.copy(order="F")
    orc.SetPressureBCs(B, 0.3, 0.0)
    assert np.array_equal(B, npo.pressure_bc(A))


@pytest.mark.parametrize("grid", GRIDS, ids=lambda g: "x".join(map(str, g)))
def test_flux_residual_update_numpy_vs_c(orc, grid):
    rng = np.random.default_rng(3)
    sh = shapes(*grid)
    A, R = rand(sh["E"], rng), rand(sh["c"], rng)
    kx, ky, kz = positive(sh["Fx"], rng), positive(sh["Fy"], rng), positive(sh["Fz"], rng)
    qx, qy, qz = F(sh["Fx"]), F(sh["Fy"]), F(sh["Fz"])
    orc.ComputeFlux(qx, qy, qz, kx, ky, kz, A, 17.0, 33.0, 9.0)
    qn = npo.ComputeFlux(kx, ky, kz, A, 17.0, 33.0, 9.0)
    for a, b in zip((qx, qy, qz), qn):
        assert np.array_equal(a, b)
    Fo = F(sh["c"])
    orc.ResidualDiffusion(Fo, R, qx, qy, qz, 17.0, 33.0, 9.0)
    assert np.array_equal(Fo, npo.ResidualDiffusion(R, *qn, 17.0, 33.0, 9.0))
    orc.ResidualDiffusionTransient(Fo, A, R, qx, qy, qz, 41.0, 17.0, 33.0, 9.0)
    assert np.array_equal(Fo, npo.ResidualDiffusionTransient(A, R, *qn, 41.0, 17.0, 33.0, 9.0))
    G, D, rc = F(sh["c"]), F(sh["c"]), positive(sh["c"], rng)
    orc.GershgorinPoisson(G, D, rc, kx, ky, kz, 1.0, 3e-3, 0.1, 0.07, 0.13)
    Gn, Dn = npo.GershgorinPoisson(rc, kx, ky, kz, 1.0, 3e-3, 0.1, 0.07, 0.13)
    assert np.array_equal(G, Gn) and np.array_equal(D, Dn)
    F0, X, Fr = rand(sh["c"], rng), rand(sh["E"], rng), rand(sh["c"], rng)
    F0n, Xn = npo.DYRELUpdate(F0, X, Fr, 0.97, 0.01, 0.3)
    orc.DYRELUpdate(F0, X, Fr, 0.97, 0.01, 0.3)
    assert np.array_equal(F0, F0n) and np.array_equal(X, Xn)
    Ty = rand(sh["Fy"], rng)
    V = [F(sh[k]) for k in ("Fx", "Fy", "Fz")]
    orc.Init_vel(*V, kx, ky, kz, A, Ty, 500.0, 17.0, 33.0, 9.0)
    for a, b in zip(V, npo.Init_vel(kx, ky, kz, A, Ty, 500.0, 17.0, 33.0, 9.0)):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("grid", GRIDS, ids=lambda g: "x".join(map(str, g)))
@pytest.mark.parametrize("version", [1, 2])
def test_weno_advection_numpy_vs_c(orc, grid, version):
    rng = np.random.default_rng(4 + version)
    sh = shapes(*grid)
    T = rand(sh["E"], rng)
    V = [rand(sh[k], rng) for k in ("Fx", "Fy", "Fz")]
    V[0].ravel(order="F")[::4] = 0.0
    Tc = T.copy(order="F")
    if version == 1:
        orc.AdvectWithWeno5(Tc, *V, 2e-3, 17.0, 33.0, 9.0, 0.0, 1.0)
    else:
        orc.AdvectWithWeno5_v2(Tc, F(sh["X6"]), F(sh["c"]), *V, 2e-3, 17.0, 33.0, 9.0, 0.0, 1.0)
    Tn = npo.advect_weno5(T, *V, 2e-3, 17.0, 33.0, 9.0, 0.0, 1.0, version)
    assert np.array_equal(Tc[1:-1, 1:-1, 1:-1], Tn[1:-1, 1:-1, 1:-1])


def _problem(grid, transient, seed, sigma=0.5):
    rng = np.random.default_rng(seed)
    sh = shapes(*grid)
    k = [fa(10.0 ** (sigma * rng.standard_normal(sh[s]))) for s in ("Fx", "Fy", "Fz")]
    X = rand(sh["E"], rng)
    R = rand(sh["c"], rng, -50, 50)
    if not transient:
        R = fa(R - R.mean())
    return X, R, k


BC = {0: ((0,) * 6, (0.0,) * 6), 1: ((0, 0, 1, 1, 0, 0), (0.0, 0.0, 1.0, 0.0, 0.0, 0.0))}


@pytest.mark.parametrize("transient", [0, 1], ids=["darcy", "heat"])
@pytest.mark.parametrize("sched", [dict(nitout=1000, ncheck=0, norm=0), dict(nitout=50, ncheck=10, norm=1)],
                         ids=["reference", "production"])
def test_dyrel_solve_numpy_vs_c(orc, transient, sched):
    grid = (6, 5, 7)
    X, R, (kx, ky, kz) = _problem(grid, transient, 7)
    h = dict(dx=0.2, dy=0.15, dz=0.1, dt=4e-3)
    bct, bcv = BC[transient]
    Xc = X.copy(order="F")
    p = orc.params(nitmax=4000, tol=1e-10, **sched)
    rc, _ = orc.dyrel_solve(Xc, R, kx, ky, kz, transient=transient, rhoC=float(transient), bc_type=bct, bc_val=bcv,
                            p=p, **h)
    Xn, rn, _ = npo.dyrel_solve(X, R, kx, ky, kz, transient=transient, rhoC=float(transient), bc_type=bct, bc_val=bcv,
                                nitmax=4000, tol=1e-10, **sched, **h)
    assert rc["status"] == rn["status"] == 0
    assert rc["iters"] == rn["iters"]
    assert rc["lmin"] == rn["lmin"]
    assert np.array_equal(Xc, Xn)


@pytest.mark.parametrize("dt_rule", [0, 1, 2], ids=["v18", "v18_constant", "v7"])
def test_hrl_steps_numpy_vs_c(orc, dt_rule):
    """Three physical steps under each time step rule (v18 variable, v18
    constant with advection sub-steps, v7), C oracle vs NumPy bit for bit."""
    from hydrothermal3d_amd import inputs
    w = inputs.WORKLOADS["hrl256"].scaled(8, 6, 8)
    T0 = inputs.initial_temperature(w)
    kfv = np.ones((9, 7, 9), order="F")
    orc.lognormal_field(kfv, *inputs.field_args(w))
    m = orc.HRLModel(8, 6, 8, w.Lx, w.Ly, w.Lz, w.Ra, w.cfl)
    m.field("Tc_ex")[...] = T0
    m.field("kfv")[...] = kfv
    p = orc.params(nitmax=20000, nitout=100, ncheck=25, norm=1, tol=1e-10)
    m.set_params(p, p, 1)
    n = npo.HRL(8, 6, 8, w.Lx, w.Ly, w.Lz, w.Ra, w.cfl)
    n.T = T0.copy()
    n.kfv = kfv.copy()
    n.prm = dict(nitmax=20000, nitout=100, ncheck=25, norm=1, tol=1e-10)
    m.set_dt_rule(dt_rule)
    n.dt_rule = dt_rule
    if dt_rule == 1:
        m.set_dt(2e-3)
        n.dt = 2e-3
    for _ in range(3):
        rc, rn = m.step(), n.step()
        assert rc["n_adv"] == rn["n_adv"]
        assert rc["thermal"]["iters"] == rn["thermal"]["iters"]
        assert rc["darcy"]["iters"] == rn["darcy"]["iters"]
        assert rc["Nu"] == rn["Nu"]
        assert rc["dt_next"] == rn["dt_next"]
    assert np.array_equal(m.field("Tc_ex")[1:-1, 1:-1, 1:-1], n.T[1:-1, 1:-1, 1:-1])
    assert np.array_equal(m.field("Pc_ex")[1:-1, 1:-1, 1:-1], n.P[1:-1, 1:-1, 1:-1])


# ---------------------------------------------------------------- closed forms
def test_conduction_fixed_point_and_nusselt(orc):
    """T = 1 - y with Dirichlet ghosts (Tbot=1, Ttop=0) and Neumann sides: the
    heat flux is exactly k dT/Ly = 1 and div q = 0, so Nu (Alban_v6:346) = 1 on
    a unit cross-section; a DYREL solve from that state does not move it."""
    nx, ny, nz = 6, 10, 5
    dy = 1.0 / ny
    yce = (np.arange(ny + 2) - 0.5) * dy
    T = fa(np.broadcast_to((1.0 - yce)[None, :, None], (nx + 2, ny + 2, nz + 2)).copy())
    orc.SetTemperatureBCs(T, 1.0, 0.0)
    sh = shapes(nx, ny, nz)
    k = [np.ones(sh[s], order="F") for s in ("Fx", "Fy", "Fz")]
    q = [F(sh[s]) for s in ("Fx", "Fy", "Fz")]
    orc.ComputeFlux(*q, *k, T, 1.0 / (1.0 / nx), 1.0 / dy, 1.0 / (1.0 / nz))
    assert np.allclose(q[1], 1.0, rtol=0, atol=1e-12)
    assert np.allclose(q[0], 0.0, atol=1e-14) and np.allclose(q[2], 0.0, atol=1e-14)
    Fr = F(sh["c"])
    orc.ResidualDiffusion(Fr, F(sh["c"]), *q, nx, 1.0 / dy, nz)
    assert np.max(np.abs(Fr)) < 1e-9
    nu = q[1][:, -1, :].sum() / (1.0 * 1.0) / (nx * nz)
    assert abs(nu - 1.0) < 1e-12
    # a whole heat solve leaves the conductive profile in place
    dt = 1e-2
    Rt = fa(-(1.0 / dt) * T[1:-1, 1:-1, 1:-1])
    X = T.copy(order="F")
    rep, work = orc.dyrel_solve(X, Rt, *k, dx=1.0 / nx, dy=dy, dz=1.0 / nz, dt=dt, transient=1, rhoC=1.0,
                                bc_type=BC[1][0], bc_val=BC[1][1], p=orc.params(tol=1e-10))
    assert rep["status"] == 0 and rep["iters"] <= 2
    assert np.max(np.abs(X[1:-1, 1:-1, 1:-1] - T[1:-1, 1:-1, 1:-1])) < 1e-12


@pytest.mark.parametrize("rhoC,transient", [(1.0, 1.0), (0.0, 0.0)])
def test_gershgorin_uniform(orc, rhoC, transient):
    """notebooks/Gershgorin_TH3D.ipynb: uniform k and h give
    G = (12/h^2 + rhoC/dt) / (6/h^2 + rhoC/dt); Darcy (rhoC = 0) gives G = 2."""
    n, h, dt = (5, 4, 6), 0.125, 3e-3
    sh = shapes(*n)
    G, D = F(sh["c"]), F(sh["c"])
    k = [np.ones(sh[s], order="F") for s in ("Fx", "Fy", "Fz")]
    orc.GershgorinPoisson(G, D, np.full(sh["c"], rhoC, order="F"), *k, transient, dt, h, h, h)
    want_D = 6.0 / h**2 + rhoC * transient / dt
    want_G = (12.0 / h**2 + rhoC * transient / dt) / want_D
    assert np.allclose(D, want_D, rtol=1e-15) and np.allclose(G, want_G, rtol=1e-15)
    if rhoC == 0.0:
        assert np.all(G == 2.0)


def test_rayleigh_quotient_on_eigenvector(orc):
    """RayleighQuotient{Numerator,Denominator}! (kernels_HT3D_dimensional.jl:349-363)
    return the eigenvalue when the update direction is an eigenvector of the
    operator: cos(pi m (i+1/2)/n) in x is an eigenvector of the Neumann
    Laplacian with eigenvalue (2 - 2 cos(pi m / n)) / h^2."""
    nx, ny, nz, m = 12, 4, 5, 3
    h = 0.1
    sh = shapes(nx, ny, nz)
    i = np.arange(nx)
    v = fa(np.broadcast_to(np.cos(np.pi * m * (i + 0.5) / nx)[:, None, None], sh["c"]).copy())
    A = F(sh["E"])
    A[1:-1, 1:-1, 1:-1] = v
    orc.SetPressureBCs(A, 0.0, 0.0)
    k = [np.ones(sh[s], order="F") for s in ("Fx", "Fy", "Fz")]
    q = [F(sh[s]) for s in ("Fx", "Fy", "Fz")]
    orc.ComputeFlux(*q, *k, A, 1 / h, 1 / h, 1 / h)
    Av = F(sh["c"])
    orc.ResidualDiffusion(Av, F(sh["c"]), *q, 1 / h, 1 / h, 1 / h)   # = -Laplace(v) = lambda v
    dtau = 0.37
    num, den = F(sh["c"]), F(sh["c"])
    orc.RayleighQuotientNumerator(num, v, F(sh["c"]), fa(-dtau * Av), dtau)
    orc.RayleighQuotientDenominator(den, v, dtau)
    lam = abs(orc.reduce(num, orc.SUM)) / abs(orc.reduce(den, orc.SUM))
    assert abs(lam - (2 - 2 * math.cos(math.pi * m / nx)) / h**2) < 1e-10 * lam


def test_weno5_closed_forms(orc):
    """dFdx_Weno5!: a constant field has zero derivative (weights -> 0.1,0.6,0.3);
    on a linear field all v_i are equal and the result is exact."""
    assert orc.weno5_js(0.0, 0.0, 0.0, 0.0, 0.0) == 0.0
    for s in (1.0, -3.25, 1e-5, 7.5e3):
        assert abs(orc.weno5_js(s, s, s, s, s) - s) <= 1e-15 * abs(s)


def test_weno5_order_of_accuracy():
    """On a smooth profile the upwind-biased WENO5 derivative converges at ~5th order."""
    errs = []
    for n in (40, 80):
        x = (np.arange(n) + 0.5) / n
        u = np.sin(2 * np.pi * x)[:, None, None] * np.ones((1, 3, 3))
        dm, _ = npo.weno_derivatives(u, 0, float(n), 2, 0.0, 2, 0.0)   # periodic
        exact = 2 * np.pi * np.cos(2 * np.pi * x)
        errs.append(np.max(np.abs(dm[:, 1, 1] - exact)))
    assert math.log2(errs[0] / errs[1]) > 4.0


@pytest.mark.parametrize("version", [1, 2])
def test_weno_constant_field_invariant(orc, version):
    """Advecting T = const (with Dirichlet values equal to the constant) leaves T unchanged."""
    sh = shapes(7, 6, 5)
    rng = np.random.default_rng(9)
    T = np.full(sh["E"], 0.625, order="F")
    V = [rand(sh[k], rng) for k in ("Fx", "Fy", "Fz")]
    if version == 1:
        orc.AdvectWithWeno5(T, *V, 1e-2, 7.0, 6.0, 5.0, 0.625, 0.625)
    else:
        orc.AdvectWithWeno5_v2(T, F(sh["X6"]), F(sh["c"]), *V, 1e-2, 7.0, 6.0, 5.0, 0.625, 0.625)
    assert np.all(T[1:-1, 1:-1, 1:-1] == 0.625)


def test_weno_v2_minus_stencil_quirk():
    """The v2 minus-side fifth difference X[I+2]-X[I-1] (Weno5_Routines_v2.jl:69)
    is kept as in the reference: on a smooth profile it makes the minus-side
    derivative measurably less accurate than v1's X[I+2]-X[I+1]
    (Weno5_Routines.jl:76), while the plus side is unchanged."""
    n = 40
    x = (np.arange(n) + 0.5) / n
    u = np.sin(2 * np.pi * x)[:, None, None] * np.ones((1, 3, 3))
    exact = 2 * np.pi * np.cos(2 * np.pi * x)
    dm1, dp1 = npo.weno_derivatives(u, 0, float(n), 2, 0.0, 2, 0.0)
    dm2, dp2 = npo.weno_derivatives(u, 0, float(n), 2, 0.0, 2, 0.0, v2_quirk=True)
    assert np.array_equal(dp1, dp2)
    e1 = np.max(np.abs(dm1[:, 1, 1] - exact))
    e2 = np.max(np.abs(dm2[:, 1, 1] - exact))
    assert e2 > 10 * e1


def test_reference_advection_known_answer(orc):
    """tools/TestAdvection.jl:33-117: a Gaussian (sigma 0.1, centre (-0.5,-0.5))
    translated at V=(1,1,0) with CFL 0.5 for 50 steps on 122x122x3 cells over
    [-1,1]^2 x [0,1] (v2 routines, Ttop=Tbot=1): the peak lands at x0 + t V."""
    ncx, ncy, ncz = 122, 122, 3
    xmin, xmax, ymin, ymax, zmin, zmax = -1.0, 1.0, -1.0, 1.0, 0.0, 1.0
    dx, dy, dz = (xmax - xmin) / ncx, (ymax - ymin) / ncy, (zmax - zmin) / ncz
    xce = np.linspace(xmin - dx / 2, xmax + dx / 2, ncx + 2)
    yce = np.linspace(ymin - dy / 2, ymax + dy / 2, ncy + 2)
    x0, y0, s = -0.5, -0.5, 0.1
    T = fa(np.broadcast_to((5 * np.exp(-((xce[:, None] - x0) ** 2) / s**2 - ((yce[None, :] - y0) ** 2) / s**2) + 1.0)
                           [:, :, None], (ncx + 2, ncy + 2, ncz + 2)).copy())
    sh = shapes(ncx, ncy, ncz)
    Vx, Vy, Vz = np.ones(sh["Fx"], order="F"), np.ones(sh["Fy"], order="F"), F(sh["Fz"])
    dt = 0.5 * min(dx, dy, dz) / 1.0
    X6, Told = F(sh["X6"]), F(sh["c"])
    t = 0.0
    for _ in range(50):
        t += dt
        orc.AdvectWithWeno5_v2(T, X6, Told, Vx, Vy, Vz, dt, 1 / dx, 1 / dy, 1 / dz, 1.0, 1.0)
    i, j = np.unravel_index(np.argmax(T[:, :, 1]), T[:, :, 1].shape)
    assert abs(xce[i] - (x0 + t)) <= dx and abs(yce[j] - (y0 + t)) <= dy
    assert np.isfinite(T).all()


# ---------------------------------------------------------------- brute force
def _assemble(apply, n):
    f0 = apply(np.zeros(n))
    M = np.column_stack([apply(np.eye(n)[:, c]) - f0 for c in range(n)])
    return M, f0


@pytest.mark.parametrize("transient", [0, 1], ids=["darcy", "heat"])
def test_dyrel_matches_dense_solve(orc, transient):
    """The converged DYREL state equals numpy.linalg.solve of the assembled
    7-point system (Neumann Darcy: the least-norm solution, compared after
    removing the mean)."""
    grid = (5, 4, 6)
    nx, ny, nz = grid
    X, R, k = _problem(grid, transient, 21)
    h = (0.2, 0.25, 1 / 6)
    dt = 0.05
    bct, bcv = BC[transient]
    N = nx * ny * nz

    def apply(x):
        A = F(shapes(*grid)["E"])
        A[1:-1, 1:-1, 1:-1] = x.reshape(grid, order="F")
        A = npo.fill_ghosts(A, bct, bcv)
        q = npo.ComputeFlux(*k, A, 1 / h[0], 1 / h[1], 1 / h[2])
        if transient:
            r = npo.ResidualDiffusionTransient(A, R, *q, 1 / dt, 1 / h[0], 1 / h[1], 1 / h[2])
        else:
            r = npo.ResidualDiffusion(R, *q, 1 / h[0], 1 / h[1], 1 / h[2])
        return r.ravel(order="F")

    M, f0 = _assemble(apply, N)
    if transient:
        x = np.linalg.solve(M, -f0)
    else:
        x = np.linalg.lstsq(M, -f0, rcond=None)[0]
    Xs = X.copy(order="F")
    rep, _ = orc.dyrel_solve(Xs, R, *k, dx=h[0], dy=h[1], dz=h[2], dt=dt, transient=transient,
                             rhoC=float(transient), bc_type=bct, bc_val=bcv,
                             p=orc.params(nitmax=200000, nitout=20, ncheck=10, norm=1, tol=1e-13))
    assert rep["status"] == 0
    got = Xs[1:-1, 1:-1, 1:-1].ravel(order="F")
    if not transient:
        got, x = got - got.mean(), x - x.mean()
    assert np.max(np.abs(got - x)) <= 1e-9 * np.max(np.abs(x))


# ---------------------------------------------------------------- invariants
def test_rp_compatibility(orc):
    """With KillNorthSouthFluxes! the buoyancy source sums to zero, the
    compatibility condition of the all-Neumann Darcy problem."""
    rng = np.random.default_rng(4)
    sh = shapes(9, 7, 8)
    T, kv = rand(sh["E"], rng), positive(sh["V"], rng)
    Ty, kx, ky, kz, Rp = F(sh["Fy"]), F(sh["Fx"]), F(sh["Fy"]), F(sh["Fz"]), F(sh["c"])
    orc.InitDarcy(Ty, kx, ky, kz, T, kv, 1.0)
    orc.KillNorthSouthFluxes(Ty)
    orc.Compute_Rp(Rp, ky, Ty, 500.0, 7.0)
    assert abs(orc.reduce(Rp, orc.SUM)) <= 1e-12 * orc.reduce(Rp, orc.SUM_ABS)


def test_operator_symmetry(orc):
    """Pairwise flux antisymmetry makes the Neumann operator symmetric: u.(A v) = v.(A u)."""
    grid = (6, 5, 4)
    rng = np.random.default_rng(8)
    sh = shapes(*grid)
    k = [positive(sh[s], rng) for s in ("Fx", "Fy", "Fz")]

    def A(x):
        E = F(sh["E"])
        E[1:-1, 1:-1, 1:-1] = x
        E = npo.pressure_bc(E)
        return npo.ResidualDiffusion(np.zeros(sh["c"]), *npo.ComputeFlux(*k, E, 3.0, 4.0, 5.0), 3.0, 4.0, 5.0)

    u, v = rng.standard_normal(sh["c"]), rng.standard_normal(sh["c"])
    a, b = np.sum(u * A(v)), np.sum(v * A(u))
    assert abs(a - b) <= 1e-12 * max(abs(a), 1.0)


def test_mean_pressure_preserved(orc):
    """With uniform k (uniform D) and a compatible source, pseudo-iterations on
    the singular Neumann problem keep mean(P) fixed."""
    X, R, k = _problem((6, 6, 6), 0, 3, sigma=0.0)
    m0 = X[1:-1, 1:-1, 1:-1].mean()
    orc.dyrel_solve(X, R, *k, dx=0.1, dy=0.1, dz=0.1, dt=1.0, transient=0, rhoC=0.0,
                    p=orc.params(nitmax=300, tol=0.0))
    assert abs(X[1:-1, 1:-1, 1:-1].mean() - m0) < 1e-12


# ---------------------------------------------------------------- failure handling (CPU)
def test_status_codes(orc):
    X, R, k = _problem((6, 6, 6), 0, 11)
    kw = dict(dx=0.1, dy=0.1, dz=0.1, dt=1.0, transient=0, rhoC=0.0)
    rep, _ = orc.dyrel_solve(X.copy(order="F"), R, *k, **kw, p=orc.params(nitmax=5, tol=0.0))
    assert rep["status"] == 10 and rep["iters"] == 5            # nitmax (silent in the reference)
    Rn = R.copy(order="F")
    Rn[2, 2, 2] = np.nan
    rep, _ = orc.dyrel_solve(X.copy(order="F"), Rn, *k, **kw, p=orc.params(nitmax=5))
    assert rep["status"] == 11                                   # v18:464 "Nan"
    rep, _ = orc.dyrel_solve(X.copy(order="F"), R, *k, **kw, p=orc.params(nitmax=2000, nitout=1, CFL=3.0))
    assert rep["status"] == 12                                   # v18:465 "Diverged"


# ---------------------------------------------------------------- textbook physics
def _onset_growth(orc, Ra, dt, n=16, nz=3, t_end=0.1, eps=1e-3):
    """Growth rate of the x-y roll (cos(pi x) sin(pi y)) of the conductive state
    in a unit box, from the closed-box model steps (thermal DYREL solve, Darcy
    DYREL solve, velocity, WENO5 advection) at a constant dt."""
    o = orc.HRLModel(n, n, nz, 1.0, 1.0, nz / n, Ra, 0.5)
    o.set_dt_rule(1)
    o.set_dt(dt)
    h = 1.0 / n
    c = (np.arange(n + 2) - 0.5) * h
    X, Y = np.meshgrid(c, c, indexing="ij")
    T = (1 - Y) + eps * np.cos(np.pi * X) * np.sin(np.pi * Y)
    o.field("Tc_ex")[...] = np.asfortranarray(np.repeat(T[:, :, None], nz + 2, axis=2))
    p = orc.params(100000, 100, 25, 1, 1e-11)
    o.set_params(p, p, 1)
    base = (1 - Y)[1:-1, 1:-1, None]
    steps = int(round(t_end / dt))
    amp = []
    for _ in range(steps):
        o.step()
        amp.append(np.abs(o.field("Tc_ex")[1:-1, 1:-1, 1:-1] - base).max())
    o.close()
    k0 = steps // 4
    return math.log(amp[-1] / amp[k0]) / (dt * (steps - 1 - k0))


@pytest.mark.parametrize("Ra", [30.0, 60.0])
def test_convection_onset_growth_rate(orc, Ra):
    """Horton-Rogers-Lapwood onset (the textbook case for the coupled model):
    a perturbation k = m = pi of the conductive state grows at
    sigma = Ra k^2/(k^2+m^2) - (k^2+m^2) = Ra/2 - 2 pi^2, so it decays below
    Ra_c = 4 pi^2 and grows above.  The steps are first order in time, so the
    rate is extrapolated from dt and dt/2 (Richardson); it must match linear
    theory within 5 % (16 cells per unit length)."""
    threads = orc.set_threads(1)     # a 768-cell grid: threading costs more than it saves
    try:
        s1, s2 = _onset_growth(orc, Ra, 0.005), _onset_growth(orc, Ra, 0.0025)
    finally:
        orc.set_threads(threads)
    sigma = 2 * s2 - s1
    theory = Ra / 2 - 2 * math.pi ** 2
    assert math.copysign(1, sigma) == math.copysign(1, theory)
    assert abs(sigma - theory) <= 0.05 * abs(theory), (sigma, theory)
