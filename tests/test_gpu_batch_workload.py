"""The 96-problem batching workload: grid edges 32-128, cubes and half-height
boxes, Ra log-uniform in [50, 2000], one lognormal permeability per problem
(hydrothermal3d_amd.inputs.batch_workloads, seeds 13000 + p).

* A sample of its problems stepped as one batch against the CPU oracle's model
  (the unfused reference kernel sequence, Dimensionless/HydroThermal3D_Alban_v6.jl
  step order): FP64 within 1e-10 relative in T, P (mean removed) and Nu with the
  same iteration counts; FP32 sweeps + FP64 defect correction within 1e-6.
* The whole workload stepped as one batch against each model stepped alone on
  the GPU: equal iteration counts, fields within 1e-10; the defect-correction
  batch within 1e-6 of the FP64 one.
"""
import numpy as np
import pytest

from helpers import rel_err

pytestmark = pytest.mark.gpu

N_PROBLEMS = 96
STEPS = 2
# every edge (32, 48, 64, 96, 128), both shapes for 32/48/64 (n^3 and n x n/2 x n),
# and the problems of smallest (p 49, Ra 52) and largest (p 22, Ra 1840) Rayleigh number
SAMPLE = (5, 14, 2, 11, 0, 18, 8, 21, 49, 22)
FP64_TOL, DC_TOL = 1e-10, 1e-6


def _params(ht, tol=1e-9):
    return ht.DyrelParams(nitmax=20000, nitout=100, ncheck=25, norm=1, tol=tol)


def _models(ht, ws, precision, tol=1e-9):
    from hydrothermal3d_amd import inputs
    ms = []
    for w in ws:
        m = ht.HydroThermal3D(w.ncx, w.ncy, w.ncz, w.Lx, w.Ly, w.Lz, w.Ra, w.cfl)
        m["Tc_ex"].set(inputs.initial_temperature(w))
        inputs.permeability_device(w, m["kfv"])
        m.set_params(_params(ht, tol), _params(ht, tol), 1, precision)
        ms.append(m)
    return ms


def _state(m):
    I = (slice(1, -1),) * 3
    P = m["Pc_ex"].get()[I]
    return {"T": m["Tc_ex"].get()[I], "P": P - P.mean()}


@pytest.fixture(scope="module")
def sample(ht, orc):
    """The sampled workloads, their device permeabilities and the oracle's
    reports and final fields after STEPS steps."""
    from hydrothermal3d_amd import inputs
    ws = inputs.batch_workloads(N_PROBLEMS, seed=13000, steps=STEPS)
    picked = [ws[p] for p in SAMPLE]
    assert {w.ncx for w in picked} == {32, 48, 64, 96, 128}
    assert {w.ncy == w.ncx for w in picked} == {True, False}
    ras = [w.Ra for w in ws]
    assert picked[-2].Ra == min(ras) and picked[-1].Ra == max(ras)
    out = []
    p = _params(ht)
    po = orc.params(p.nitmax, p.nitout, p.ncheck, p.norm, p.tol)
    for w in picked:
        g = ht.HydroThermal3D(w.ncx, w.ncy, w.ncz, w.Lx, w.Ly, w.Lz, w.Ra, w.cfl)
        inputs.permeability_device(w, g["kfv"])
        kfv = g["kfv"].get()
        g.close()
        o = orc.HRLModel(w.ncx, w.ncy, w.ncz, w.Lx, w.Ly, w.Lz, w.Ra, w.cfl)
        o.field("Tc_ex")[...] = inputs.initial_temperature(w)
        o.field("kfv")[...] = kfv
        o.set_params(po, po, 1)
        reps = [o.step() for _ in range(STEPS)]
        I = (slice(1, -1),) * 3
        P = o.field("Pc_ex")[I].copy()
        out.append({"reps": reps, "T": o.field("Tc_ex")[I].copy(), "P": P - P.mean()})
        o.close()
    return picked, out


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "fp32dc"])
def test_batch_sample_matches_oracle(ht, sample, precision):
    ws, ref = sample
    ms = _models(ht, ws, precision)
    b = ht.HydroThermal3DBatch(ms, _params(ht), _params(ht), precision=precision, graphs=True)
    reps = [b.step() for _ in range(STEPS)]
    tol = FP64_TOL if precision == 0 else DC_TOL
    for q, (w, m, o) in enumerate(zip(ws, ms, ref)):
        name = f"problem {SAMPLE[q]} ({w.ncx}x{w.ncy}x{w.ncz}, Ra {w.Ra:.0f})"
        for s in range(STEPS):
            a, c = reps[s][q], o["reps"][s]
            assert abs(a["Nu"] - c["Nu"]) <= tol * abs(c["Nu"]), (name, s, a["Nu"], c["Nu"])
            if precision == 0:
                for kind in ("thermal", "darcy"):
                    assert abs(a[kind]["iters"] - c[kind]["iters"]) <= 1, (name, s, kind)
        got = _state(m)
        for k in ("T", "P"):
            e = rel_err(got[k], o[k])
            assert e <= tol, (name, k, e)
    b.close()
    for m in ms:
        m.close()


def test_batch_96_problem_workload(ht):
    """One step of the whole workload, three ways on the same inputs, at the
    configurations' tolerance 1e-8: each model stepped alone in FP64, the FP64
    batch and the defect-correction batch (same iteration counts within one
    check interval)."""
    from hydrothermal3d_amd import inputs
    tol = 1e-8
    ws = inputs.batch_workloads(N_PROBLEMS, seed=13000, steps=1)
    assert sum(w.ncell for w in ws) == 36784128
    single = _models(ht, ws, 0, tol)
    rs = [m.step() for m in single]
    fs = [_state(m) for m in single]
    for m in single:
        m.close()
    res = {}
    for precision in (0, 1):
        ms = _models(ht, ws, precision, tol)
        b = ht.HydroThermal3DBatch(ms, _params(ht, tol), _params(ht, tol), precision=precision, graphs=True)
        res[precision] = (b.step(), [_state(m) for m in ms], b.stats())
        b.close()
        for m in ms:
            m.close()
    rb, fb, st = res[0]
    assert st["repacks"] > 0 and st["graph_launches"] > 0
    for q in range(N_PROBLEMS):
        for kind in ("thermal", "darcy"):
            assert rb[q][kind]["iters"] == rs[q][kind]["iters"], (q, kind)
            assert rb[q][kind]["status"] == rs[q][kind]["status"] == 0, (q, kind)
        assert abs(rb[q]["Nu"] - rs[q]["Nu"]) <= FP64_TOL * abs(rs[q]["Nu"]), q
        for k in ("T", "P"):
            assert rel_err(fb[q][k], fs[q][k]) <= FP64_TOL, (q, k)
    rd, fd, _ = res[1]
    for q in range(N_PROBLEMS):
        assert rd[q]["thermal"]["status"] == rd[q]["darcy"]["status"] == 0, q
        for kind in ("thermal", "darcy"):
            assert abs(rd[q][kind]["iters"] - rb[q][kind]["iters"]) <= _params(ht).ncheck, (q, kind)
        assert abs(rd[q]["Nu"] - rb[q]["Nu"]) <= DC_TOL * abs(rb[q]["Nu"]), q
        for k in ("T", "P"):
            assert rel_err(fd[q][k], fb[q][k]) <= DC_TOL, (q, k, rel_err(fd[q][k], fb[q][k]))
