"""FP32 inner sweeps + FP64 defect correction against the plain FP64 GPU solve.

No reference counterpart exists (the reference is FP64 only); the bar is the
one set for the mixed-precision form: <= 1e-6 relative error in T, P (gauge
fixed) and Nu against the FP64 GPU path on the same inputs."""
import numpy as np
import pytest

from helpers import shapes, rand, rel_err

pytestmark = pytest.mark.gpu


def _problem(grid, transient, seed):
    rng = np.random.default_rng(seed)
    sh = shapes(*grid)
    k = [np.asfortranarray(10.0 ** (0.5 * rng.standard_normal(sh[s]))) for s in ("Fx", "Fy", "Fz")]
    X = rand(sh["E"], rng)
    R = rand(sh["c"], rng, -50, 50)
    if not transient:
        R = np.asfortranarray(R - R.mean())
    return X, R, k


@pytest.mark.parametrize("transient", [0, 1], ids=["darcy", "heat"])
@pytest.mark.parametrize("grid", [(33, 17, 25), (64, 32, 48)], ids=lambda g: "x".join(map(str, g)))
def test_dc_solve_matches_fp64(ht, grid, transient):
    X, R, k = _problem(grid, transient, 17)
    h = dict(dx=1.0 / grid[0], dy=0.5 / grid[1], dz=1.0 / grid[2], dt=1e-3)
    bc = dict(bc_type=(0, 0, 1, 1, 0, 0), bc_val=(0.0, 0.0, 1.0, 0.0, 0.0, 0.0)) if transient else {}
    prm = ht.production_params(tol=1e-11, nitout=100, ncheck=25)
    out, reps = {}, {}
    for prec in (0, 1):
        D = [ht.DeviceArray.from_host(a) for a in (X, R, *k)]
        rep = ht.dyrel_solve(*D, transient=transient, rhoC=float(transient), params=prm, precision=prec, **h, **bc)
        assert rep["status"] == 0, rep
        out[prec] = D[0].get()[1:-1, 1:-1, 1:-1]
        reps[prec] = rep
    # the anchors test the iterates the FP64 check iterations test, and the
    # counts include the final update: equal up to a decision one check apart
    assert abs(reps[1]["iters"] - reps[0]["iters"]) <= prm.ncheck, (reps[0]["iters"], reps[1]["iters"])
    assert reps[1]["iters"] % prm.ncheck == 0, reps[1]["iters"]
    a, b = out[1], out[0]
    if not transient:
        a, b = a - a.mean(), b - b.mean()
    assert rel_err(a, b) <= 1e-6


def test_hrl_steps_dc_matches_fp64(ht):
    from hydrothermal3d_amd import inputs
    w = inputs.WORKLOADS["hrl256"].scaled(48, 24, 40)
    res = {}
    for prec in (0, 1):
        m = ht.HydroThermal3D(w.ncx, w.ncy, w.ncz, w.Lx, w.Ly, w.Lz, w.Ra, w.cfl)
        m["Tc_ex"].set(inputs.initial_temperature(w))
        inputs.permeability_device(w, m["kfv"])
        p = ht.production_params(tol=1e-11)
        m.set_params(p, p, 1, prec)
        nus = [m.step()["Nu"] for _ in range(3)]
        res[prec] = (m["Tc_ex"].get()[1:-1, 1:-1, 1:-1], m["Pc_ex"].get()[1:-1, 1:-1, 1:-1], np.array(nus))
        m.close()
    (T0, P0, N0), (T1, P1, N1) = res[0], res[1]
    assert rel_err(T1, T0) <= 1e-6
    assert rel_err(P1 - P1.mean(), P0 - P0.mean()) <= 1e-6
    assert np.max(np.abs(N1 - N0) / np.abs(N0)) <= 1e-6


@pytest.mark.parametrize("name", ["hrl512_fault", "hrl768_fault"])
def test_fault_copy_oracle_and_dc(ht, orc, name):
    """Scaled copies of the faulted workloads (lognormal k x 100 in a 60-degree
    slab, and for the larger one mu = exp(-ln10 T)): the FP64 GPU steps against
    the CPU oracle (1e-10, iterations +-1), the FP32+DC steps against the FP64
    GPU steps (1e-6 in T, P and Nu)."""
    from hydrothermal3d_amd import inputs
    w = inputs.WORKLOADS[name].scaled(64, 32, 64)
    p = ht.production_params(tol=1e-9)
    res = {}
    for prec in (0, 1):
        m = ht.HydroThermal3D(w.ncx, w.ncy, w.ncz, w.Lx, w.Ly, w.Lz, w.Ra, w.cfl)
        m["Tc_ex"].set(inputs.initial_temperature(w))
        inputs.permeability_device(w, m["kfv"])
        m.set_viscosity(w.mu_a)
        m.set_params(p, p, 1, prec)
        kfv = m["kfv"].get()
        reps = [m.step() for _ in range(2)]
        res[prec] = (m["Tc_ex"].get()[1:-1, 1:-1, 1:-1], m["Pc_ex"].get()[1:-1, 1:-1, 1:-1],
                     np.array([r["Nu"] for r in reps]), reps)
        m.close()
    o = orc.HRLModel(w.ncx, w.ncy, w.ncz, w.Lx, w.Ly, w.Lz, w.Ra, w.cfl)
    o.set_viscosity(w.mu_a)
    o.field("Tc_ex")[...] = inputs.initial_temperature(w)
    o.field("kfv")[...] = kfv
    po = orc.params(p.nitmax, p.nitout, p.ncheck, p.norm, p.tol)
    o.set_params(po, po, 1)
    ro = [o.step() for _ in range(2)]
    T0, P0, N0, r0 = res[0]
    for a, c in zip(r0, ro):
        assert abs(a["thermal"]["iters"] - c["thermal"]["iters"]) <= 1
        assert abs(a["darcy"]["iters"] - c["darcy"]["iters"]) <= 1
    To, Po = o.field("Tc_ex")[1:-1, 1:-1, 1:-1], o.field("Pc_ex")[1:-1, 1:-1, 1:-1]
    assert rel_err(T0, To) <= 1e-10
    assert rel_err(P0 - P0.mean(), Po - Po.mean()) <= 1e-10
    assert np.max(np.abs(N0 - np.array([r["Nu"] for r in ro])) / np.abs(N0)) <= 1e-10
    o.close()
    T1, P1, N1, _ = res[1]
    assert rel_err(T1, T0) <= 1e-6
    assert rel_err(P1 - P1.mean(), P0 - P0.mean()) <= 1e-6
    assert np.max(np.abs(N1 - N0) / np.abs(N0)) <= 1e-6


def _fault_steps(ht, w, p, prec, nsteps):
    from hydrothermal3d_amd import inputs
    m = ht.HydroThermal3D(w.ncx, w.ncy, w.ncz, w.Lx, w.Ly, w.Lz, w.Ra, w.cfl)
    m["Tc_ex"].set(inputs.initial_temperature(w))
    inputs.permeability_device(w, m["kfv"])
    m.set_viscosity(w.mu_a)
    m.set_params(p, p, 1, prec)
    reps = [m.step() for _ in range(nsteps)]
    T, P = m["Tc_ex"].get()[1:-1, 1:-1, 1:-1], m["Pc_ex"].get()[1:-1, 1:-1, 1:-1]
    m.close()
    return T, P - P.mean(), np.array([r["Nu"] for r in reps]), reps


def test_768_fault_copy_dc_at_tol_1e8(ht):
    """The 768^3-class faulted configuration (Ra 2000, lognormal k x 100 in a
    60-degree slab, mu = exp(-ln10 T)) on a 128x64x128 copy at its tolerance
    1e-8, two steps: the FP32+DC steps against the FP64 steps, 1e-6 in T, P
    and Nu, with the same iteration counts.  Both solves end on the same kind
    of state (the tested iterate plus one update, ghosts from the tested
    iterate); T at the cells where the fault meets the bottom wall moves by up
    to 1e-5 when the ghosts come from the final iterate instead, since the
    closed wall then sees a pressure gradient across it.  Both forms are also
    compared with the converged solution (FP64 at tol 1e-11): DC is no further
    from it than FP64 is, within 1e-6."""
    from hydrothermal3d_amd import inputs
    w = inputs.WORKLOADS["hrl768_fault"].scaled(128, 64, 128)
    T0, P0, N0, r0 = _fault_steps(ht, w, ht.production_params(tol=1e-8), 0, 2)
    T1, P1, N1, r1 = _fault_steps(ht, w, ht.production_params(tol=1e-8), 1, 2)
    Tr, Pr, Nr, _ = _fault_steps(ht, w, ht.production_params(tol=1e-11), 0, 2)
    for a, c in zip(r0, r1):
        assert a["thermal"]["status"] == c["thermal"]["status"] == 0
        assert a["darcy"]["status"] == c["darcy"]["status"] == 0
        assert a["thermal"]["iters"] == c["thermal"]["iters"]
        assert a["darcy"]["iters"] == c["darcy"]["iters"]
    assert rel_err(T1, T0) <= 1e-6
    assert rel_err(P1, P0) <= 1e-6
    assert np.max(np.abs(N1 - N0) / np.abs(N0)) <= 1e-6
    assert rel_err(T1, Tr) <= rel_err(T0, Tr) + 1e-6
    assert rel_err(P1, Pr) <= rel_err(P0, Pr) + 1e-6


@pytest.mark.parametrize("ncheck", [1, 2, 3, 7])
@pytest.mark.parametrize("transient", [0, 1], ids=["darcy", "heat"])
def test_dc_anchor_intervals(ht, transient, ncheck):
    """Anchor intervals of 1, 2, 3 and 7 sweeps.  The first sweep of each run
    is queued before the anchor norm is read back (a two-sweep launch when
    pairs are on and the run has room for one) and its result is dropped when
    the norm shows convergence; with 1 or 2 sweeps per anchor it is the whole
    run.  Pairs forced and pairs off must give the same solve bit for bit, and
    both must stay within 1e-6 of the FP64 solve."""
    grid = (40, 21, 30)
    X, R, k = _problem(grid, transient, 29)
    h = dict(dx=1.0 / grid[0], dy=0.5 / grid[1], dz=1.0 / grid[2], dt=1e-3)
    bc = dict(bc_type=(0, 0, 1, 1, 0, 0), bc_val=(0.0, 0.0, 1.0, 0.0, 0.0, 0.0)) if transient else {}
    prm = ht.production_params(tol=1e-11, nitout=100, ncheck=ncheck)
    out = {}
    try:
        for key, prec, pairs in (("fp64", 0, 1), ("dc_single", 1, 0), ("dc_pairs", 1, 2)):
            ht._lib.call("ht3d_tune_pairs", pairs)
            D = [ht.DeviceArray.from_host(a) for a in (X, R, *k)]
            rep = ht.dyrel_solve(*D, transient=transient, rhoC=float(transient), params=prm, precision=prec,
                                 **h, **bc)
            assert rep["status"] == 0, (key, rep)
            out[key] = (rep, D[0].get()[1:-1, 1:-1, 1:-1])
    finally:
        ht._lib.call("ht3d_tune_pairs", 1)
    (r0, x0), (r1, x1), (r2, x2) = out["fp64"], out["dc_single"], out["dc_pairs"]
    assert np.array_equal(x1, x2)
    for key in ("iters", "n_checks", "n_lambda", "lmin", "nF_abs", "nF_rel"):
        assert r1[key] == r2[key], (key, r1[key], r2[key])
    assert r1["iters"] % ncheck == 0 or r1["iters"] == prm.nitmax
    a, b = x1, x0
    if not transient:
        a, b = a - a.mean(), b - b.mean()
    assert rel_err(a, b) <= 1e-6


@pytest.mark.parametrize("pairs", [0, 2], ids=["single", "pairs"])
def test_dc_status_nitmax_and_nan(ht, pairs):
    """The defect-correction solve stops at nitmax with status 10 and exactly
    nitmax pseudo-iterations, also when nitmax is not a multiple of the anchor
    interval (a last run shorter than the interval, possibly queued whole
    before the anchor norm is read), and NaN data raises like the FP64 solve
    (HydroThermal3D_v18.jl:464)."""
    grid = (16, 9, 12)
    X, R, k = _problem(grid, 0, 5)
    h = dict(dx=0.1, dy=0.1, dz=0.1, dt=1.0)
    ht._lib.call("ht3d_tune_pairs", pairs)
    try:
        for nitmax, ncheck in ((3, 50), (7, 3), (50, 7), (41, 4)):
            D = [ht.DeviceArray.from_host(a) for a in (X, R, *k)]
            prm = ht.DyrelParams(nitmax=nitmax, nitout=1000, ncheck=ncheck, tol=0.0)
            rep = ht.dyrel_solve(*D, transient=0, rhoC=0.0, params=prm, precision=1, **h)
            assert rep["status"] == 10 and rep["iters"] == nitmax, (nitmax, ncheck, rep)
        Rn = np.asfortranarray(R.copy())
        Rn[3, 3, 3] = np.nan
        D = [ht.DeviceArray.from_host(a) for a in (X, Rn, *k)]
        with pytest.raises(ht.SolveError, match="Nan"):
            ht.dyrel_solve(*D, transient=0, rhoC=0.0, params=ht.DyrelParams(nitmax=50, nitout=1000, ncheck=7),
                           precision=1, **h)
    finally:
        ht._lib.call("ht3d_tune_pairs", 1)
