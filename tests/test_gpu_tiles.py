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


# ------------------------------------------------ face coefficients from a vertex field
KV_TILES = (14, 15, 16, 17, 18, 19)   # the two-cell tiles: the vertex-coefficient kernels run on these


def _kv_problem(grid, transient, seed=29):
    """Face arrays that are the face averages of a lognormal vertex field, as
    InitThermal!/InitDarcy! form them (kernels_HT3D.jl:82-107): the sum of the
    four vertex values left to right, then times 0.25."""
    rng = np.random.default_rng(seed)
    sh = shapes(*grid)
    kv = np.asfortranarray(10.0 ** (0.5 * rng.standard_normal(sh["V"])))
    kx = np.asfortranarray((kv[:, :-1, :-1] + kv[:, 1:, :-1] + kv[:, :-1, 1:] + kv[:, 1:, 1:]) * 0.25)
    ky = np.asfortranarray((kv[:-1, :, :-1] + kv[1:, :, :-1] + kv[:-1, :, 1:] + kv[1:, :, 1:]) * 0.25)
    kz = np.asfortranarray((kv[:-1, :-1, :] + kv[1:, :-1, :] + kv[:-1, 1:, :] + kv[1:, 1:, :]) * 0.25)
    X = rand(sh["E"], rng)
    R = rand(sh["c"], rng, -50, 50)
    if not transient:
        R = np.asfortranarray(R - R.mean())
    return X, R, [kx, ky, kz], kv


def _kv_solve(ht, grid, transient, tile, variant, with_kv, precision=0):
    X, R, k, kv = _kv_problem(grid, transient)
    if tile is None:
        ht._lib.call("ht3d_tune_set", 1, -1, 0)
    else:
        ht._lib.call("ht3d_tune_set", 0, tile, variant)
    try:
        D = [ht.DeviceArray.from_host(a) for a in (X, R, *k)]
        dkv = ht.DeviceArray.from_host(kv) if with_kv else None
        h = dict(dx=1.0 / grid[0], dy=0.5 / grid[1], dz=1.3 / grid[2], dt=1e-3)
        bc = dict(bc_type=(0, 0, 1, 1, 0, 0), bc_val=(0.0, 0.0, 1.0, 0.0, 0.0, 0.0)) if transient else {}
        prm = ht.DyrelParams(nitmax=20000, nitout=100, ncheck=25, norm=1, tol=1e-10)
        rep = ht.dyrel_solve(*D, transient=transient, rhoC=float(transient), params=prm, kv=dkv,
                             precision=precision, **h, **bc)
        ty, kz, xc, var = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        ht._lib.call("ht3d_tune_last_tile", C.byref(ty), C.byref(kz), C.byref(xc), C.byref(var))
        return rep, D[0].get(), (ty.value, kz.value, xc.value, var.value)
    finally:
        ht._lib.call("ht3d_tune_set", 1, -1, 0)


@pytest.mark.parametrize("transient", [0, 1], ids=["darcy", "heat"])
@pytest.mark.parametrize("grid", [(70, 37, 45), (16, 9, 70), (260, 7, 9), (133, 20, 19)],
                         ids=lambda g: "x".join(map(str, g)))
def test_vertex_coefficients_bitwise(ht, grid, transient):
    """The kernels that form kx, ky, kz from the vertex field (variant bit 8)
    give the face-array kernel's fields and reports bit for bit, on every
    two-cell tile, with and without the stored diagonal (bit 1) and the
    non-temporal store (bit 2).  On an odd x extent, or without the vertex
    field, the bit is dropped and the face arrays are read."""
    r0, X0, cfg0 = _kv_solve(ht, grid, transient, REF_TILE, REF_VARIANT, False)
    assert cfg0 == (4, 4, 1, REF_VARIANT)
    for n, tile in enumerate(KV_TILES):
        # bit 1: stored diagonal; bit 2: non-temporal store; bit 4 (rolled z march) is FP32 only
        variant = 8 | (2 if n % 2 else 0) | (4 if n % 3 == 1 else 0) | (1 if n % 3 == 0 else 0)
        r, X, cfg = _kv_solve(ht, grid, transient, tile, variant, True)
        if grid[0] % 2 == 0:
            assert cfg[2] == 2 and cfg[3] == variant & ~4, (tile, cfg)
        else:
            assert cfg[2] == 1 and not cfg[3] & 8, (tile, cfg)
        for key in ("iters", "status", "n_checks", "n_lambda", "lmin", "h1", "h2", "nF_abs", "nF_rel"):
            assert r0[key] == r[key], (tile, variant, key, r0[key], r[key])
        assert np.array_equal(X0, X), (tile, variant)
    # pinned to the vertex bit but without the vertex field: the face arrays are read
    r, X, cfg = _kv_solve(ht, grid, transient, KV_TILES[0], 8, False)
    assert not cfg[3] & 8
    assert np.array_equal(X0, X)


def test_vertex_coefficients_tuned(ht):
    """With the vertex field the tuner also times the vertex kernels; whatever it
    picks, the solve equals the face-array solve bit for bit."""
    grid = (96, 48, 64)
    r0, X0, _ = _kv_solve(ht, grid, 0, None, 0, False)
    r1, X1, cfg = _kv_solve(ht, grid, 0, None, 0, True)
    for key in ("iters", "status", "n_checks", "n_lambda", "lmin", "h1", "h2", "nF_abs", "nF_rel"):
        assert r0[key] == r1[key], (key, cfg)
    assert np.array_equal(X0, X1), cfg


KV_TILES_F32 = KV_TILES + (20, 21, 22, 23, 24)   # FP32 also has them on the four-cell tiles


@pytest.mark.parametrize("transient", [0, 1], ids=["darcy", "heat"])
@pytest.mark.parametrize("grid", [(70, 37, 45), (260, 7, 9), (136, 20, 19)], ids=lambda g: "x".join(map(str, g)))
def test_vertex_coefficients_dc_bitwise(ht, grid, transient):
    """Defect correction with the vertex field: the FP32 faces are the FP32
    averages of the FP32 vertex field, so the face-array FP32 kernels, the FP32
    vertex kernels (two and four cells per lane, all variants) and whatever
    the tuner picks (pairs and band pairs included, and the FP64 anchor
    residual from the vertex field) give the same solve bit for bit; and it
    stays within 1e-6 of the FP64 solve."""
    r0, X0, cfg0 = _kv_solve(ht, grid, transient, REF_TILE, REF_VARIANT, True, precision=1)
    assert cfg0 == (4, 4, 1, REF_VARIANT)
    for n, tile in enumerate(KV_TILES_F32):
        variant = 8 | (2 if n % 2 else 0) | (4 if n % 3 == 1 else 0)
        r, X, cfg = _kv_solve(ht, grid, transient, tile, variant, True, precision=1)
        assert cfg[3] & 8, (tile, cfg)
        for key in ("iters", "status", "n_checks", "n_lambda", "dc_outer", "nF_abs", "nF_rel"):
            assert r0[key] == r[key], (tile, variant, key, r0[key], r[key])
        assert np.array_equal(X0, X), (tile, variant)
    rt, Xt, cfgt = _kv_solve(ht, grid, transient, None, 0, True, precision=1)
    for key in ("iters", "status", "n_checks", "n_lambda", "dc_outer", "nF_abs", "nF_rel"):
        assert r0[key] == rt[key], (key, cfgt, r0[key], rt[key])
    assert np.array_equal(X0, Xt), cfgt
    r64, X64, _ = _kv_solve(ht, grid, transient, None, 0, True, precision=0)
    m = np.zeros(X64.shape, bool)
    m[1:-1, 1:-1, 1:-1] = True
    err = np.max(np.abs(X0[m] - X64[m])) / np.max(np.abs(X64[m]))
    assert err <= 1e-6, err


BAND_TILES = {11: 4, 12: 4, 13: 4, 14: 2}   # band pair tiles with 2 or 4 cells per lane: cells per lane


@pytest.mark.parametrize("transient", [0, 1], ids=["darcy", "heat"])
@pytest.mark.parametrize("grid", [(128, 37, 45), (256, 20, 19)], ids=lambda g: "x".join(map(str, g)))
def test_vertex_coefficients_band_pairs_bitwise(ht, grid, transient):
    """The FP32 band pairs of the defect correction with the coefficients formed
    from the vertex field (pair variant bit 8) against the same tiles on the
    FP32 face arrays and against single launches only: identical solves."""
    r0, X0, _ = _kv_solve_pairs(ht, grid, transient, None)
    for tile, cpl in BAND_TILES.items():
        for variant in (0, 8):
            r, X, (xc, var) = _kv_solve_pairs(ht, grid, transient, (tile, variant))
            assert (xc, var) == (cpl, variant), (tile, xc, var)
            for key in ("iters", "status", "n_checks", "n_lambda", "dc_outer", "nF_abs", "nF_rel"):
                assert r0[key] == r[key], (tile, variant, key, r0[key], r[key])
            assert np.array_equal(X0, X), (tile, variant)


def _kv_solve_pairs(ht, grid, transient, pair):
    """DC solve with the vertex field: single launches only (pair None) or the
    band pair tile/variant pinned; returns the report, X and the pair tile's
    (cells per lane, variant)."""
    X, R, k, kv = _kv_problem(grid, transient)
    if pair is None:
        ht._lib.call("ht3d_tune_pairs", 0)
    else:
        ht._lib.call("ht3d_tune_force_pair", *pair)
    try:
        D = [ht.DeviceArray.from_host(a) for a in (X, R, *k)]
        dkv = ht.DeviceArray.from_host(kv)
        h = dict(dx=1.0 / grid[0], dy=0.5 / grid[1], dz=1.3 / grid[2], dt=1e-3)
        bc = dict(bc_type=(0, 0, 1, 1, 0, 0), bc_val=(0.0, 0.0, 1.0, 0.0, 0.0, 0.0)) if transient else {}
        prm = ht.DyrelParams(nitmax=20000, nitout=100, ncheck=25, norm=1, tol=1e-10)
        rep = ht.dyrel_solve(*D, transient=transient, rhoC=float(transient), params=prm, kv=dkv, precision=1,
                             **h, **bc)
        ty, kz, xc, var = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        ht._lib.call("ht3d_tune_last_pair_tile", C.byref(ty), C.byref(kz), C.byref(xc), C.byref(var))
        return rep, D[0].get(), (xc.value, var.value)
    finally:
        ht._lib.call("ht3d_tune_pairs", 1)
        ht._lib.call("ht3d_tune_force_pair", -1, 0)
