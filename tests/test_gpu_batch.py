"""The batched step (ht3d_batch_*) against the same models stepped one by one.

Each problem keeps its own grid, Rayleigh number, permeability, viscosity,
time step rule and WENO version; the batch shares the DYREL parameters.  The
batched solves run the same pseudo-iteration body and the same double-double
reductions as the single-problem solves, so iteration counts must agree and
fields agree to round-off (bar: 1e-10 relative, the FP64 parity bar).  Graph
replay and kernel-by-kernel launches must give identical results.
"""
import math

import numpy as np
import pytest

from helpers import rel_err

pytestmark = pytest.mark.gpu

SIZES = [(16, 8, 16), (24, 12, 20), (20, 20, 20), (13, 9, 11), (32, 16, 24), (12, 12, 12)]
TOL = 1e-10


def _workloads(sizes=SIZES):
    from hydrothermal3d_amd import inputs
    ws = inputs.batch_workloads(len(sizes), seed=13000, steps=3)
    return [inputs.replace(w, ncx=a, ncy=b, ncz=c, Lx=1.0, Ly=b / a, Lz=c / a, modes=64)
            for w, (a, b, c) in zip(ws, sizes)]


def _models(ht, ws, p, precision, setup=None):
    from hydrothermal3d_amd import inputs
    ms = []
    for q, w in enumerate(ws):
        m = ht.HydroThermal3D(w.ncx, w.ncy, w.ncz, w.Lx, w.Ly, w.Lz, w.Ra, w.cfl)
        m["Tc_ex"].set(inputs.initial_temperature(w))
        inputs.permeability_device(w, m["kfv"])
        m.set_params(p, p, 2 if q == 5 else 1, precision)
        if q == 1:
            m.set_viscosity(math.log(10.0))
        if q == 3:
            m.set_dt_rule(ht._lib.DT_V7)
        if setup is not None:
            setup(q, m)
        ms.append(m)
    return ms


def _fields(m):
    I = (slice(1, -1),) * 3
    P = m["Pc_ex"].get()[I]
    return {"T": m["Tc_ex"].get()[I], "P": P - P.mean(), "Vx": m["Vx"].get(), "Vy": m["Vy"].get(),
            "Vz": m["Vz"].get()}


def _run(ht, precision, graphs, nsteps=3, sizes=SIZES, setup=None):
    ws = _workloads(sizes)
    p = ht.DyrelParams(nitmax=20000, nitout=100, ncheck=25, norm=1, tol=1e-9)
    ms = _models(ht, ws, p, precision, setup)
    b = ht.HydroThermal3DBatch(ms, p, p, precision=precision, graphs=graphs)
    reps = [b.step() for _ in range(nsteps)]
    stats = b.stats()
    out = [_fields(m) for m in ms]
    b.close()
    for m in ms:
        m.close()
    return reps, out, stats


def _run_single(ht, precision, nsteps=3, setup=None):
    ws = _workloads()
    p = ht.DyrelParams(nitmax=20000, nitout=100, ncheck=25, norm=1, tol=1e-9)
    ms = _models(ht, ws, p, precision, setup)
    reps = [[m.step() for m in ms] for _ in range(nsteps)]
    out = [_fields(m) for m in ms]
    for m in ms:
        m.close()
    return reps, out


def _compare_with_single(rb, fb, rs, fs, precision):
    for s, (stepb, steps_) in enumerate(zip(rb, rs)):
        for q, (a, c) in enumerate(zip(stepb, steps_)):
            for kind in ("thermal", "darcy"):
                assert a[kind]["iters"] == c[kind]["iters"], f"step {s} problem {q} {kind}"
                assert a[kind]["status"] == c[kind]["status"]
                assert a[kind]["n_lambda"] == c[kind]["n_lambda"]
            assert a["n_adv"] == c["n_adv"]
            assert abs(a["Nu"] - c["Nu"]) <= TOL * abs(c["Nu"]), f"step {s} problem {q}"
            assert abs(a["dt_next"] - c["dt_next"]) <= TOL * c["dt_next"]
            if precision == 1:
                assert a["dc_outer_darcy"] == c["dc_outer_darcy"]
    for q, (a, c) in enumerate(zip(fb, fs)):
        for k in a:
            assert rel_err(a[k], c[k]) <= TOL, f"problem {q} field {k}: {rel_err(a[k], c[k])}"


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "fp32dc"])
def test_batch_matches_single_models(ht, precision):
    rb, fb, stats = _run(ht, precision, graphs=True)
    rs, fs = _run_single(ht, precision)
    _compare_with_single(rb, fb, rs, fs, precision)
    assert stats["graphs"] > 0 and stats["graph_launches"] >= stats["graphs"]
    assert stats["repacks"] > 0          # problems of different sizes stop at different checks


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "fp32dc"])
def test_batch_graphs_equal_eager(ht, precision):
    rg, fg, sg = _run(ht, precision, graphs=True, nsteps=2)
    re_, fe, se = _run(ht, precision, graphs=False, nsteps=2)
    assert se["graph_launches"] == 0
    for a, c in zip(fg, fe):
        for k in a:
            assert np.array_equal(a[k], c[k]), k
    for stepg, stepe in zip(rg, re_):
        for a, c in zip(stepg, stepe):
            assert a["thermal"] == c["thermal"] and a["darcy"] == c["darcy"] and a["Nu"] == c["Nu"]


def test_batch_advection_substeps(ht):
    """The constant-dt rule on every model: the batch advects every model's
    first sub-step in one launch per pass (WENO v1 and v2 in the same launch)
    and the further sub-steps per model; the results are those of the models
    stepped alone, with more than one sub-step in some steps."""
    def setup(q, m):
        m.set_dt_rule(ht._lib.DT_V18_CONSTANT)
        m.set_dt(2e-4)
    # Two steps: with dt fixed, max|V| of the third problem grows about 15x a
    # step (7.3e3, 1.1e5), alone as in the batch, and a third step would need
    # more than the 1e6 sub-steps the step accepts.
    rb, fb, _ = _run(ht, 0, graphs=True, nsteps=2, setup=setup)
    rs, fs = _run_single(ht, 0, nsteps=2, setup=setup)
    _compare_with_single(rb, fb, rs, fs, 0)
    assert max(r["n_adv"] for step in rb for r in step) > 1
    for stepb, steps_ in zip(rb, rs):
        for a, c in zip(stepb, steps_):
            assert a["dt_next"] == c["dt_next"] == 2e-4


def test_batch_nan_raises(ht):
    ws = _workloads()[:3]
    p = ht.DyrelParams(nitmax=2000, nitout=100, ncheck=25, norm=1, tol=1e-9)
    ms = _models(ht, ws, p, 0)
    T = ms[1]["Tc_ex"].get()
    T[3, 3, 3] = np.nan
    ms[1]["Tc_ex"].set(T)
    b = ht.HydroThermal3DBatch(ms, p, p)
    with pytest.raises(ht.SolveError, match="Nan"):
        b.step()
    b.close()
    for m in ms:
        m.close()


# every ncx a multiple of 4 (the 2- and 4-cell batch tiles run), 136 leaving a
# partial x tile for both
EVEN = [(16, 8, 16), (24, 12, 20), (36, 20, 12), (136, 8, 6), (12, 12, 12)]
BATCH_TILE_CPL = (1, 1, 1, 1, 1, 1, 2, 2, 2, 4, 4, 4, 2, 2, 2, 4, 2, 2, 2, 2, 4, 4, 4)


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "fp32dc"])
def test_batch_tiles_bitwise(ht, precision):
    """Each batch tile pinned in turn (ht3d_batch_tune_force).  The tiles with 2
    or 4 x-cells per lane compute every cell with the one-cell operations in
    the same order, so steps, reports and fields equal the default tile's bit
    for bit.  4 cells per lane is not offered to FP64 sweeps: pinned there, the
    batch falls back to the default tile."""
    n = ht._lib.lib().ht3d_batch_tile_count()
    assert n == len(BATCH_TILE_CPL)
    ref = None
    try:
        for t in [1] + [t for t in range(n) if t != 1]:
            ht._lib.call("ht3d_batch_tune_force", t)
            reps, out, stats = _run(ht, precision, graphs=True, nsteps=2, sizes=EVEN)
            cpl = BATCH_TILE_CPL[t] if (precision == 1 or BATCH_TILE_CPL[t] <= 2) else 1
            used = stats["tile_cpl"][2:] if precision == 1 else stats["tile_cpl"][:2]
            assert used == [cpl, cpl], (t, stats["tile_cpl"])
            if ref is None:
                ref = (reps, out)
                continue
            for stepr, stepo in zip(reps, ref[0]):
                for a, c in zip(stepr, stepo):
                    for kind in ("thermal", "darcy"):
                        for key in ("iters", "status", "n_lambda", "n_checks", "lmin", "nF_abs"):
                            assert a[kind][key] == c[kind][key], (t, kind, key)
                    assert a["Nu"] == c["Nu"] and a["dt_next"] == c["dt_next"], t
            for a, c in zip(out, ref[1]):
                for k in a:
                    assert np.array_equal(a[k], c[k]), (t, k)
    finally:
        ht._lib.call("ht3d_batch_tune_force", -1)
    with pytest.raises(ht._lib.HT3DError):
        ht._lib.call("ht3d_batch_tune_force", n)


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "fp32dc"])
def test_batch_grid_modes_bitwise(ht, precision):
    """The persistent grid and the one-block-per-item grid walk the same items
    with the same operations: identical reports and fields."""
    res = []
    try:
        for mode in (0, 1):
            ht._lib.call("ht3d_batch_tune_force", 7 if precision == 0 else 10)
            ht._lib.call("ht3d_batch_tune_force_grid", mode)
            reps, out, stats = _run(ht, precision, graphs=True, nsteps=2, sizes=EVEN)
            assert stats["grid_mode"][2 * precision:2 * precision + 2] == [mode, mode]
            res.append((reps, out))
    finally:
        ht._lib.call("ht3d_batch_tune_force", -1)
        ht._lib.call("ht3d_batch_tune_force_grid", -1)
    (ra, fa), (rb, fb) = res
    for sa, sb in zip(ra, rb):
        for a, c in zip(sa, sb):
            assert a["thermal"] == c["thermal"] and a["darcy"] == c["darcy"] and a["Nu"] == c["Nu"]
    for a, c in zip(fa, fb):
        for k in a:
            assert np.array_equal(a[k], c[k]), k


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "fp32dc"])
def test_batch_vertex_bitwise(ht, precision):
    """The vertex-coefficient form of the batched launches (face coefficients
    formed from each problem's vertex field; the problem with a viscosity keeps
    its face arrays for the Darcy solve) against the face-array form, on a
    two- and a four-cell tile: identical reports and fields."""
    res = []
    tile = 7 if precision == 0 else 10   # 4 x 8 planes, 2 (FP64) or 4 (FP32) cells per lane
    try:
        for mode in (0, 1):
            ht._lib.call("ht3d_batch_tune_force", tile)
            ht._lib.call("ht3d_batch_tune_vertex", mode)
            reps, out, stats = _run(ht, precision, graphs=True, nsteps=2, sizes=EVEN)
            assert stats["vertex"][2 * precision:2 * precision + 2] == [mode, mode], stats["vertex"]
            res.append((reps, out))
    finally:
        ht._lib.call("ht3d_batch_tune_force", -1)
        ht._lib.call("ht3d_batch_tune_vertex", -1)
    (ra, fa), (rb, fb) = res
    for sa, sb in zip(ra, rb):
        for a, c in zip(sa, sb):
            assert a["thermal"] == c["thermal"] and a["darcy"] == c["darcy"] and a["Nu"] == c["Nu"]
    for a, c in zip(fa, fb):
        for k in a:
            assert np.array_equal(a[k], c[k]), k
    with pytest.raises(ht._lib.HT3DError):
        ht._lib.call("ht3d_batch_tune_vertex", 2)


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "fp32dc"])
def test_batch_pairs_bitwise(ht, precision):
    """Two pseudo-iterations per launch on bands of whole rows, one, two or
    (FP32) four x-cells per lane (each pair tile pinned in turn,
    ht3d_batch_tune_force_pair) against single launches only
    (ht3d_batch_tune_pairs(0)): the pair kernel computes every cell with the
    single iteration's operations, and runs of pairs come in even numbers, so
    steps, reports and fields are identical bit for bit."""
    import ctypes as C
    lib = ht._lib.lib()
    n = lib.ht3d_batch_pair_tile_count()
    cpl = []
    for t in range(n):
        r, z, c = C.c_int(), C.c_int(), C.c_int()
        ht._lib.call("ht3d_batch_pair_tile_info", t, C.byref(r), C.byref(z), C.byref(c))
        cpl.append(c.value)
    tiles = [t for t in range(n) if precision == 1 or cpl[t] < 4]   # 4 cells per lane: FP32 sweeps only
    try:
        ht._lib.call("ht3d_batch_tune_pairs", 0)
        ref_reps, ref_out, ref_stats = _run(ht, precision, graphs=True, nsteps=2, sizes=EVEN)
        assert ref_stats["pair_rows"] == [0, 0, 0, 0]
        ht._lib.call("ht3d_batch_tune_pairs", 1)
        seen = set()
        for t in tiles:
            ht._lib.call("ht3d_batch_tune_force_pair", t)
            for graphs in (True, False):
                reps, out, stats = _run(ht, precision, graphs=graphs, nsteps=2, sizes=EVEN)
                used = stats["pair_rows"][2 * precision:2 * precision + 2]
                assert min(used) > 0, (t, stats["pair_rows"])
                seen.add((tuple(used), tuple(stats["pair_z"][2 * precision:2 * precision + 2]),
                          tuple(stats["pair_cpl"][2 * precision:2 * precision + 2])))
                for stepr, stepo in zip(reps, ref_reps):
                    for a, c in zip(stepr, stepo):
                        assert a["thermal"] == c["thermal"] and a["darcy"] == c["darcy"], (t, graphs)
                        assert a["Nu"] == c["Nu"] and a["dt_next"] == c["dt_next"], (t, graphs)
                for a, c in zip(out, ref_out):
                    for k in a:
                        assert np.array_equal(a[k], c[k]), (t, graphs, k)
        assert len(seen) == len(tiles)
    finally:
        ht._lib.call("ht3d_batch_tune_pairs", -1)
        ht._lib.call("ht3d_batch_tune_force_pair", -1)
    for bad in (n,):
        with pytest.raises(ht._lib.HT3DError):
            ht._lib.call("ht3d_batch_tune_force_pair", bad)
    with pytest.raises(ht._lib.HT3DError):
        ht._lib.call("ht3d_batch_tune_pairs", 2)
