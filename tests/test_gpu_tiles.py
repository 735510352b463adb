"""Tile configurations of the fused iteration kernel against one another.

Every configuration performs the same per-cell operations in the same order;
only the grouping of cells into blocks and lanes differs.  So whole solves with
a pinned configuration must give bit-identical fields and reports to the
one-cell-per-lane reference configuration (4x4, variant 0): in FP64 and in the
FP32 sweeps of the defect-correction solve, both operators, on grids that leave
partial tiles in every direction.  A tile with 2 or 4 cells per lane needs an
x extent divisible by that number; otherwise the pinned tile runs as its twin
with the most cells per lane the extent allows, which the tuner report must
show.  (The check sums are double-double sums, so a
different block grouping leaves them unchanged except at an exact rounding
tie.)
"""
import ctypes as C

import numpy as np
import pytest

from helpers import shapes, rand

pytestmark = pytest.mark.gpu

REF_TILE, REF_VARIANT = 2, 0          # 4 rows x 4 planes, one cell per lane
VECTOR_TILES = {14: 2, 15: 2, 16: 2, 17: 2, 18: 2, 19: 2, 20: 4, 21: 4, 22: 4, 23: 4, 24: 4}   # tile: cells per lane


def _problem(grid, transient, seed=23):
    rng = np.random.default_rng(seed)
    sh = shapes(*grid)
    k = [np.asfortranarray(10.0 ** (0.5 * rng.standard_normal(sh[s]))) for s in ("Fx", "Fy", "Fz")]
    X = rand(sh["E"], rng)
    R = rand(sh["c"], rng, -50, 50)
    if not transient:
        R = np.asfortranarray(R - R.mean())
    return X, R, k


def _solve(ht, grid, transient, precision, tile, variant):
    X, R, k = _problem(grid, transient)
    ht._lib.call("ht3d_tune_set", 0, tile, variant)
    try:
        D = [ht.DeviceArray.from_host(a) for a in (X, R, *k)]
        h = dict(dx=1.0 / grid[0], dy=0.5 / grid[1], dz=1.3 / grid[2], dt=1e-3)
        bc = dict(bc_type=(0, 0, 1, 1, 0, 0), bc_val=(0.0, 0.0, 1.0, 0.0, 0.0, 0.0)) if transient else {}
        prm = ht.DyrelParams(nitmax=20000, nitout=100, ncheck=25, norm=1, tol=1e-10)
        rep = ht.dyrel_solve(*D, transient=transient, rhoC=float(transient), params=prm, precision=precision,
                             **h, **bc)
        ty, kz, xc, var = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        ht._lib.call("ht3d_tune_last_tile", C.byref(ty), C.byref(kz), C.byref(xc), C.byref(var))
        return rep, D[0].get(), (ty.value, kz.value, xc.value, var.value)
    finally:
        ht._lib.call("ht3d_tune_set", 1, -1, 0)


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "fp32dc"])
@pytest.mark.parametrize("transient", [0, 1], ids=["darcy", "heat"])
@pytest.mark.parametrize("grid", [(70, 37, 45), (16, 9, 70), (133, 20, 19), (260, 7, 9)],
                         ids=lambda g: "x".join(map(str, g)))
def test_cells_per_lane_bitwise(ht, grid, transient, precision):
    r0, X0, cfg0 = _solve(ht, grid, transient, precision, REF_TILE, REF_VARIANT)
    assert cfg0 == (4, 4, 1, REF_VARIANT)
    for n, (tile, cpl) in enumerate(VECTOR_TILES.items()):
        variant = (3 * n + 1) % 8          # every variant bit set and clear across the tiles
        r, X, cfg = _solve(ht, grid, transient, precision, tile, variant)
        runs = max(c for c in (1, 2, 4) if c <= cpl and grid[0] % c == 0)
        assert cfg[2] == runs, (tile, cfg)
        for key in ("iters", "status", "n_checks", "n_lambda", "lmin", "h1", "h2", "nF_abs", "nF_rel"):
            assert r0[key] == r[key], (tile, variant, key, r0[key], r[key])
        assert np.array_equal(X0, X), (tile, variant)


def test_tile_table_reports(ht):
    """The table holds 14 one-cell tiles, 6 two-cell and 5 four-cell tiles;
    pinning an index past its end is refused."""
    n = ht._lib.lib().ht3d_tile_count()
    assert n == 25
    with pytest.raises(ht._lib.HT3DError):
        ht._lib.call("ht3d_tune_set", 0, n, 0)
    ht._lib.call("ht3d_tune_set", 1, -1, 0)
