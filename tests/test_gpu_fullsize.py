"""The hot path at the headline size, 256x128x256 cells, against the CPU oracle
on identical inputs:

* one FP64 physical step of the convection workload (lognormal k, Ra=500):
  iteration counts, T, P (gauge-fixed) and Nu;
* the FP32 sweeps + FP64 defect correction against the FP64 path, within 1e-6;
* both fused v18 loops for a fixed number of pseudo-iterations on the
  synthetic v18 geometry: reports and fields bit for bit.

The oracle needs a few seconds per solve at this size on a 16-thread host."""
import numpy as np
import pytest

from helpers import rel_err
from test_gpu_v18 import _compare, _problem, _run

pytestmark = pytest.mark.gpu

GRID = (256, 128, 256)


def _gauge(P):
    inn = P[1:-1, 1:-1, 1:-1]
    return inn - inn.mean()


def _models(ht, orc, precision):
    from hydrothermal3d_amd import inputs
    w = inputs.WORKLOADS["hrl256"]
    assert (w.ncx, w.ncy, w.ncz) == GRID
    prm = ht.production_params(tol=1e-8, ncheck=50)
    T0 = inputs.initial_temperature(w)
    g = ht.HydroThermal3D(w.ncx, w.ncy, w.ncz, w.Lx, w.Ly, w.Lz, w.Ra, w.cfl)
    g["Tc_ex"].set(T0)
    inputs.permeability_device(w, g["kfv"])
    g.set_params(prm, prm, 1, precision)
    o = None
    if orc is not None:
        o = orc.HRLModel(w.ncx, w.ncy, w.ncz, w.Lx, w.Ly, w.Lz, w.Ra, w.cfl)
        o.field("Tc_ex")[...] = T0
        o.field("kfv")[...] = g["kfv"].get()
        po = orc.params(prm.nitmax, prm.nitout, prm.ncheck, prm.norm, prm.tol)
        o.set_params(po, po, 1)
    return g, o


def test_headline_step_matches_oracle(ht, orc):
    from hydrothermal3d_amd import _lib
    g, o = _models(ht, orc, _lib.FP64)
    try:
        print("full-size FP64 step vs the oracle", flush=True)
        rg, rc = g.step(), o.step()
        for kind in ("thermal", "darcy"):
            assert abs(rg[kind]["iters"] - rc[kind]["iters"]) <= 1, (kind, rg[kind], rc[kind])
        assert abs(rg["Nu"] - rc["Nu"]) <= 1e-10 * abs(rc["Nu"]), (rg["Nu"], rc["Nu"])
        I = (slice(1, -1),) * 3
        assert rel_err(g["Tc_ex"].get()[I], o.field("Tc_ex")[I]) <= 1e-10
        assert rel_err(_gauge(g["Pc_ex"].get()), _gauge(o.field("Pc_ex"))) <= 1e-10
    finally:
        g.close()
        o.close()


def test_headline_step_defect_correction(ht):
    from hydrothermal3d_amd import _lib
    out = []
    for prec in (_lib.FP64, _lib.FP32_DC):
        g, _ = _models(ht, None, prec)
        try:
            print(f"full-size step, precision {prec}", flush=True)
            r = g.step()
            out.append((r, g["Tc_ex"].get()[1:-1, 1:-1, 1:-1], _gauge(g["Pc_ex"].get())))
        finally:
            g.close()
    (r64, T64, P64), (r32, T32, P32) = out
    assert rel_err(T32, T64) <= 1e-6
    assert rel_err(P32, P64) <= 1e-6
    assert abs(r32["Nu"] - r64["Nu"]) <= 1e-6 * abs(r64["Nu"])


@pytest.mark.parametrize("kind", ["heat", "darcy"])
def test_v18_loops_full_size_bitwise(ht, orc, kind):
    """60 pseudo-iterations (checks and lambda updates at 1, 2 and 50) of the
    fused v18 loop at 256x128x256, with the immersed-surface layer."""
    d = _problem(GRID, True, 1e14)
    d["wN"][:] = np.where(d["wN"] < 1.0, 0.9, 1.0)
    print(f"full-size v18 {kind}", flush=True)
    _compare(d, _run(ht, orc, d, kind, dict(nitmax=60, nitout=50, ncheck=0, norm=0, tol=0.0)))
