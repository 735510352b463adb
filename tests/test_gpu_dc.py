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
    out = {}
    for prec in (0, 1):
        D = [ht.DeviceArray.from_host(a) for a in (X, R, *k)]
        rep = ht.dyrel_solve(*D, transient=transient, rhoC=float(transient), params=prm, precision=prec, **h, **bc)
        assert rep["status"] == 0, rep
        out[prec] = D[0].get()[1:-1, 1:-1, 1:-1]
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
