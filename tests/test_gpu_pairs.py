"""Two pseudo-iterations per launch (ht3d_tune_pairs) against one per launch.

A pair launch performs the single iteration's operations in the same order, so
whole solves with pairs forced on and with pairs off must agree bit for bit
(fields, iteration counts, lambda updates), in FP64 and in the FP32 sweeps of
the defect-correction solve, on grids that leave partial tiles in every
direction and with both boundary-condition tables.
"""
import numpy as np
import pytest

from helpers import shapes, rand

pytestmark = pytest.mark.gpu


def _problem(grid, transient, seed=11):
    rng = np.random.default_rng(seed)
    sh = shapes(*grid)
    k = [np.asfortranarray(10.0 ** (0.5 * rng.standard_normal(sh[s]))) for s in ("Fx", "Fy", "Fz")]
    X = rand(sh["E"], rng)
    R = rand(sh["c"], rng, -50, 50)
    if not transient:
        R = np.asfortranarray(R - R.mean())
    return X, R, k


def _solve(ht, grid, transient, precision, mode):
    X, R, k = _problem(grid, transient)
    ht._lib.call("ht3d_tune_pairs", mode)
    try:
        D = [ht.DeviceArray.from_host(a) for a in (X, R, *k)]
        h = dict(dx=1.0 / grid[0], dy=0.5 / grid[1], dz=1.3 / grid[2], dt=1e-3)
        bc = dict(bc_type=(0, 0, 1, 1, 0, 0), bc_val=(0.0, 0.0, 1.0, 0.0, 0.0, 0.0)) if transient else {}
        prm = ht.DyrelParams(nitmax=20000, nitout=100, ncheck=25, norm=1, tol=1e-10)
        rep = ht.dyrel_solve(*D, transient=transient, rhoC=float(transient), params=prm, precision=precision,
                             **h, **bc)
        import ctypes as C
        ty, kz, var = C.c_int(), C.c_int(), C.c_int()
        ht._lib.call("ht3d_tune_last_pair", C.byref(ty), C.byref(kz), C.byref(var))
        return rep, D[0].get(), ty.value
    finally:
        ht._lib.call("ht3d_tune_pairs", 1)


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "fp32dc"])
@pytest.mark.parametrize("transient", [0, 1], ids=["darcy", "heat"])
@pytest.mark.parametrize("grid", [(70, 37, 45), (133, 20, 19), (16, 9, 70)], ids=lambda g: "x".join(map(str, g)))
def test_pairs_bitwise(ht, grid, transient, precision):
    r0, X0, rows0 = _solve(ht, grid, transient, precision, 0)
    r2, X2, rows2 = _solve(ht, grid, transient, precision, 2)
    assert rows0 == 0 and rows2 > 0          # pairs off, then on
    for key in ("iters", "status", "n_checks", "n_lambda", "lmin", "h1", "h2", "nF_abs"):
        assert r0[key] == r2[key], key
    assert np.array_equal(X0, X2)


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "fp32dc"])
def test_sweep_direction_does_not_change_results(ht, precision):
    """Alternating the sweep direction between launches (ht3d_tune_sweep) only
    reorders the tiles of a launch: fields and reports are bit-identical."""
    out = []
    try:
        for alt in (0, 1):
            ht._lib.call("ht3d_tune_sweep", alt)
            out.append(_solve(ht, (70, 37, 45), 1, precision, 1))
    finally:
        ht._lib.call("ht3d_tune_sweep", 1)
    (r0, X0, _), (r1, X1, _) = out
    for key in ("iters", "status", "n_checks", "n_lambda", "lmin", "h1", "h2", "nF_abs"):
        assert r0[key] == r1[key], key
    assert np.array_equal(X0, X1)


PAIR_TILES = {0: 1, 1: 1, 2: 1, 3: 1, 4: 1, 5: 2, 6: 2, 7: 2, 8: 4, 9: 4, 10: 4,   # pair tile: x cells per lane
              11: 4, 12: 4, 13: 4, 14: 2, 15: 1}
BAND_ROWS = {11: 8, 12: 16, 13: 16, 14: 16, 15: 16}   # band tiles: rows of 64 lanes


def _pair_tile_runs(tile, nx, precision):
    """Whether pair tile `tile` runs on this grid: nx divisible by its cells per
    lane; in FP64 no 4-cell tile and no 2-cell tile with 16 rows; the band
    tiles (bands of whole rows) in FP32 only and when a band holds at least 3
    rows."""
    cpl = PAIR_TILES[tile]
    if tile in BAND_ROWS:
        return precision == 1 and nx % cpl == 0 and 64 * BAND_ROWS[tile] * cpl // nx >= 3
    if precision == 0 and (cpl == 4 or tile == 7):
        return False
    return nx % cpl == 0


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "fp32dc"])
@pytest.mark.parametrize("transient", [0, 1], ids=["darcy", "heat"])
@pytest.mark.parametrize("grid", [(70, 37, 45), (16, 9, 70), (260, 7, 9), (133, 20, 19)],
                         ids=lambda g: "x".join(map(str, g)))
def test_every_pair_tile_bitwise(ht, grid, transient, precision):
    """Each pair tile pinned in turn (ht3d_tune_force_pair), cycling through the
    variants: bit-identical to pairs off, or not used when the grid does not
    allow it (the report then shows no pair tile).  Tiles 11-15 are bands of
    whole rows (the batch's band-pair kernels on a one-entry table)."""
    import ctypes as C
    r0, X0, rows0 = _solve(ht, grid, transient, precision, 0)
    assert rows0 == 0
    try:
        for tile, cpl in PAIR_TILES.items():
            variant = tile % 4
            ht._lib.call("ht3d_tune_force_pair", tile, variant)
            r, X, rows = _solve(ht, grid, transient, precision, 1)
            ty, kz, xc, var = C.c_int(), C.c_int(), C.c_int(), C.c_int()
            ht._lib.call("ht3d_tune_last_pair_tile", C.byref(ty), C.byref(kz), C.byref(xc), C.byref(var))
            if _pair_tile_runs(tile, grid[0], precision):
                want_var = 0 if tile in BAND_ROWS else variant   # the band tiles have no variants
                assert (xc.value, var.value) == (cpl, want_var), (tile, xc.value, var.value)
            else:
                assert rows == 0 and xc.value == 0, (tile, rows, xc.value)
            for key in ("iters", "status", "n_checks", "n_lambda", "lmin", "h1", "h2", "nF_abs"):
                assert r0[key] == r[key], (tile, key)
            assert np.array_equal(X0, X), tile
    finally:
        ht._lib.call("ht3d_tune_force_pair", -1, 0)
