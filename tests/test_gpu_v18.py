"""Fused v18 DYREL solves (ht3d_v18_thermal_solve / ht3d_v18_darcy_solve)
against the oracle's unfused kernel sequence of HydroThermal3D_v18.jl:432-478
and :489-537 on identical inputs: the same iterations, statuses, lambda
updates and norms, and the same fields bit for bit (unknown with ghosts, air
and immersed cells; PC and G; phi*rho_f of the last Darcy iteration except at
immersed cells).  Grids leave partial tiles in every direction."""
import numpy as np
import pytest

from helpers import v18_problem

pytestmark = pytest.mark.gpu

GRIDS = [(16, 12, 6), (70, 37, 9), (37, 70, 13), (250, 106, 3)]   # the last: the v18 grid (v18:291-294)
KEYS = ("iters", "status", "n_checks", "n_lambda", "lmax", "lmin", "dtau", "h1", "h2", "nF_abs", "nF_rel", "nF_ini")


def _problem(grid, immersed, dt_s, seed=3):
    d = v18_problem(*grid, seed=seed, immersed=immersed)
    d["heat"].update(dt=dt_s / d["sc"].t, norm_scale=1e30)
    d["darcy"].update(dt=dt_s / d["sc"].t, norm_scale=1e30)
    return d


def _run(ht, orc, d, kind, prm, kv=False):
    """kv: also hand the solve the vertex field whose face averages the
    coefficients are (ktv for heat, kfv for Darcy)."""
    from hydrothermal3d_amd import dimensional as dim
    from hydrothermal3d_amd.solver import SolveError
    dev = lambda a: ht.DeviceArray.from_host(np.asfortranarray(a))

    def solve(fn, *args, **kw):   # NaN / divergence raise, as the reference does; keep the report
        try:
            return fn(*args, **kw)
        except SolveError as e:
            assert e.report is not None
            return e.report

    grid = d["T0"].shape
    PCd, Gd = ht.DeviceArray.zeros(grid), ht.DeviceArray.zeros(grid)
    p = orc.params(**prm)
    gp = ht.DyrelParams(**prm)
    if kind == "heat":
        Xh = d["T"].copy(order="F")
        rep_o, w = orc.v18_thermal_solve(Xh, d["T0"], d["rhoC"], *d["kT"], d["phc"], d["wN"], **d["heat"], p=p)
        Xd = dev(d["T"])
        rep_g = solve(dim.v18_thermal_solve, Xd, dev(d["T0"]), dev(d["rhoC"]), *map(dev, d["kT"]), dev(d["phc"]),
                                      dev(d["wN"]), **d["heat"], params=gp, PC=PCd, G=Gd,
                      kv=dev(d["ktv"]) if kv else None)
        extra = None
    else:
        Xh, ph = d["P"].copy(order="F"), d["phirho"].copy(order="F")
        rep_o, w = orc.v18_darcy_solve(Xh, d["T"], ph, d["phirhof0"], d["rhofv"], *d["kP"], d["phc"], d["wN"],
                                       **d["darcy"], p=p)
        Xd, phd = dev(d["P"]), dev(d["phirho"])
        rep_g = solve(dim.v18_darcy_solve, Xd, dev(d["T"]), phd, dev(d["phirhof0"]), dev(d["rhofv"]), *map(dev, d["kP"]),
                                    dev(d["phc"]), dev(d["wN"]), **d["darcy"], params=gp, PC=PCd, G=Gd,
                      kv=dev(d["kfv"]) if kv else None)
        extra = (phd.get(), ph)
    return rep_o, rep_g, (Xd.get(), Xh), (PCd.get(), w["PC"]), (Gd.get(), w["G"]), extra


def _compare(d, out):
    rep_o, rep_g, X, PC, G, extra = out
    for k in KEYS:
        assert rep_g[k] == rep_o[k], (k, rep_g[k], rep_o[k])
    assert np.array_equal(X[0], X[1])
    assert np.array_equal(PC[0], PC[1]) and np.array_equal(G[0], G[1])
    if extra is not None:
        keep = d["phc"][1:-1, 1:-1, 1:-1] != -1.0
        assert np.array_equal(extra[0][keep], extra[1][keep])


@pytest.mark.parametrize("kind", ["heat", "darcy"])
@pytest.mark.parametrize("immersed", [False, True], ids=["flat", "immersed"])
@pytest.mark.parametrize("grid", GRIDS, ids=lambda g: "x".join(map(str, g)))
def test_fixed_iterations_bitwise(ht, orc, grid, immersed, kind):
    """250 pseudo-iterations (checks and lambda updates at 1, 2, 100, 200)."""
    print(f"v18 {kind} {grid} immersed={immersed}", flush=True)
    d = _problem(grid, immersed, 1e14)
    if immersed:
        d["wN"][:] = np.where(d["wN"] < 1.0, 0.9, 1.0)
    _compare(d, _run(ht, orc, d, kind, dict(nitmax=250, nitout=100, ncheck=0, norm=0, tol=0.0)))


@pytest.mark.parametrize("kind", ["heat", "darcy"])
@pytest.mark.parametrize("grid", GRIDS[:2], ids=lambda g: "x".join(map(str, g)))
def test_converged_solve(ht, orc, grid, kind):
    """To relative tolerance 1e-10 with extra checks every 25 iterations."""
    print(f"v18 converge {kind} {grid}", flush=True)
    d = _problem(grid, False, 1e15, seed=8)
    out = _run(ht, orc, d, kind, dict(nitmax=20000, nitout=100, ncheck=25, norm=1, tol=1e-10))
    assert out[0]["status"] == 0
    _compare(d, out)


def test_divergence_detected_alike(ht, orc):
    """Small immersed-surface weights make the heat loop diverge (the
    extrapolation is outside the Gershgorin bound): the same iteration and
    norms on both paths."""
    d = _problem((16, 12, 6), True, 1e15)
    print("v18 divergence (16,12,6)", flush=True)
    out = _run(ht, orc, d, "heat", dict(nitmax=2000, nitout=100, ncheck=0, norm=0, tol=1e-10))
    assert out[0]["status"] == ht._lib.DIVERGED
    _compare(d, out)


V18_TWIN = {6: 0, 7: 1, 8: 3, 9: 5, 10: 0, 11: 1, 12: 3, 13: 5}   # 2-cell tile -> its 1-cell twin (odd nx)
V18_FACE_TWIN = {10: 6, 11: 7, 12: 8, 13: 9}   # vertex-field tile -> its face-array twin (no vertex field)


@pytest.mark.parametrize("kind", ["heat", "darcy"])
@pytest.mark.parametrize("grid", [(70, 37, 9), (33, 20, 7)], ids=lambda g: "x".join(map(str, g)))
def test_every_tile_bitwise(ht, orc, grid, kind):
    """Every v18 tile pinned in turn (1 and 2 x-cells per lane; the 2-cell
    tiles also forming the face coefficients from the vertex field), with
    immersed cells and with the vertex field given: bit-identical to the
    oracle.  On an odd nx a 2-cell tile runs as its 1-cell twin, and without
    the vertex field a vertex tile as its face-array twin, which the tuner
    report shows."""
    import ctypes as C
    d = _problem(grid, True, 1e14)
    d["wN"][:] = np.where(d["wN"] < 1.0, 0.9, 1.0)
    ref = None
    try:
        n = ht._lib.lib().ht3d_v18_tile_count()
        assert n == 14
        for t, kv in [(t, True) for t in range(n)] + [(t, False) for t in V18_FACE_TWIN]:
            print(f"v18 tile {t} {kind} {grid} vertex field {kv}", flush=True)
            ht._lib.call("ht3d_v18_tune", t, None)
            out = _run(ht, orc, d, kind, dict(nitmax=120, nitout=100, ncheck=0, norm=0, tol=0.0), kv=kv)
            last = C.c_int()
            ht._lib.call("ht3d_v18_tune", t, C.byref(last))
            want = V18_TWIN.get(t, t) if grid[0] % 2 else (t if kv else V18_FACE_TWIN.get(t, t))
            assert last.value == want, (t, kv, last.value)
            _compare(d, out)
            ref = out[2][0] if ref is None else ref
            assert np.array_equal(out[2][0], ref)
    finally:
        ht._lib.call("ht3d_v18_tune", -1, None)


@pytest.mark.parametrize("immersed", [False, True], ids=["flat", "immersed"])
@pytest.mark.parametrize("grid", [(24, 16, 8), (70, 30, 9), (250, 106, 3)], ids=lambda g: "x".join(map(str, g)))
def test_model_steps_match_oracle(ht, orc, grid, immersed):
    """Three passes of the v18 time loop (the steady initialisation step, the
    intrusion step, one more) by hydrothermal3d_amd.v18.V18Model against the
    oracle's kernel sequence.  The permeabilities come from exp() and 10^x,
    whose device and host results may differ by an ulp; the Rayleigh quotient
    sum(-dtau F0 (F - Fcit)) cancels nearly equal residuals and amplifies that
    (the solves themselves are bit-exact on equal inputs, above).  So the
    tolerance is measured: a second oracle run with every permeability moved by
    one ulp at random gives the spread of each quantity, and the GPU must stay
    within 10x that spread (plus 1e-12 relative; 1e-8 relative for lambda_min,
    h1 and h2; 1e-11 absolute for the relative norm, which sits at round-off
    level once converged).  Iteration
    counts, statuses and advection sub-steps agree exactly."""
    from helpers import V18OracleModel
    from hydrothermal3d_amd.v18 import V18Model, v18_geometry
    geo = v18_geometry(*grid, seed=4, immersed=immersed)
    prm = dict(nitmax=300, nitout=100, tol=1e-18)
    ref = V18OracleModel(*grid, geo, orc.params(**prm), orc.params(**prm))
    alt = V18OracleModel(*grid, geo, orc.params(**prm), orc.params(**prm), ulp_seed=9)
    gpu = V18Model(*grid, geo, ht.DyrelParams(**prm), ht.DyrelParams(**prm))

    def close(got, want, other, what, floor=1e-300):
        tol = 10.0 * abs(other - want) + 1e-12 * abs(want) + floor
        assert abs(got - want) <= tol, (what, got, want, other)

    for s in range(3):
        print(f"v18 model step {s} {grid} immersed={immersed}", flush=True)
        ro, ra, rg = ref.step(), alt.step(), gpu.step()
        assert rg["n_adv"] == ro["n_adv"]
        # device-side diagnostics (v18:575-577) against the host arrays
        assert rg["T_max"] == pytest.approx(np.max(ref.T) * ref.sc.T, rel=1e-8)
        assert rg["P_min"] == pytest.approx(np.min(ref.P) * ref.sc.sigma, rel=1e-8)
        assert rg["Vy_max"] == pytest.approx(np.max(ref.Vy) * ref.sc.V, rel=1e-8,
                                             abs=1e-8 * np.max(np.abs(ref.Vy)) * ref.sc.V)
        for solve in ("thermal", "darcy"):
            for k in ("iters", "status", "n_checks", "n_lambda"):
                assert rg[solve][k] == ro[solve][k], (s, solve, k)
            # lambda_min is a cancellation-dominated quotient: a random one-ulp
            # sample can understate a systematic exp/10^x difference by 20x on
            # it, so it also gets a 1e-8 relative floor (the iteration decisions
            # it drives are asserted exactly above)
            for k in ("lmax", "lmin", "h1", "h2"):
                close(rg[solve][k], ro[solve][k], ra[solve][k], (s, solve, k), floor=1e-8 * abs(ro[solve][k]))
            # a relative norm at round-off level (1e-13) carries no digits: absolute floor
            close(rg[solve]["nF_rel"], ro[solve]["nF_rel"], ra[solve]["nF_rel"], (s, solve), floor=1e-11)
        for name, host, host2 in (("Tc_ex", ref.T, alt.T), ("Pc_ex", ref.P, alt.P), ("Vx", ref.Vx, alt.Vx),
                                  ("Vy", ref.Vy, alt.Vy), ("Vz", ref.Vz, alt.Vz)):
            got = gpu[name].get()
            scale = np.max(np.abs(host))
            err, spread = np.max(np.abs(got - host)) / scale, np.max(np.abs(host2 - host)) / scale
            print(f"  {name}: gpu-oracle {err:.2e}, one-ulp spread {spread:.2e}", flush=True)
            assert err <= 10.0 * spread + 1e-12, (s, name, err, spread)


def test_model_output_and_restart(ht, tmp_path):
    """V18Model writes the v18 VTK fields (read back equal to output_fields) and
    a breakpoint; a model restarted from it takes the same next step as the one
    that continued (bit for bit)."""
    from hydrothermal3d_amd.v18 import V18Model, v18_geometry
    from hydrothermal3d_amd.vtkio import read_vtr
    grid = (24, 16, 8)
    geo = v18_geometry(*grid, seed=6)
    prm = ht.DyrelParams(nitmax=200, nitout=100, tol=1e-18)
    m = V18Model(*grid, geo, prm, prm)
    print("v18 output/restart", flush=True)
    m.step()
    m.step()
    fields = m.output_fields()
    path = m.save_vtk(tmp_path / "HT3DOutput00002")
    (x, y, z), f = read_vtr(path)
    assert len(x) == grid[0] and len(y) == grid[1] and len(z) == grid[2]
    assert set(f) == {"Phase", "Pressure", "Temperature", "Velocity", "kThermal", "Perm", "Density", "Viscosity",
                      "Peclet", "HeatCapacity"}
    for k, v in fields.items():
        if k == "Velocity":
            for a, b in zip(f[k], v):
                assert np.array_equal(a, b, equal_nan=True)
        else:
            assert np.array_equal(f[k], v, equal_nan=True), k
    air = geo["phc"][1:-1, 1:-1, 1:-1] == 1.0
    assert np.all(np.isnan(f["Temperature"][air])) and np.all(np.isfinite(f["Temperature"][~air]))
    bp = m.save_breakpoint(tmp_path / "Breakpoint00002")
    r = V18Model(*grid, geo, prm, prm)
    r.restart(bp, 2)
    a, b = m.step(), r.step()
    assert a["thermal"]["iters"] == b["thermal"]["iters"] and a["n_adv"] == b["n_adv"]
    assert np.array_equal(m["Tc_ex"].get(), r["Tc_ex"].get())
    assert np.array_equal(m["Pc_ex"].get(), r["Pc_ex"].get())
