"""The allocation-free DYREL solve (ht3d_dyrel_solve_noalloc) against
ht3d_dyrel_solve, both called through the C ABI on the same inputs.

The allocation-free entry takes every device buffer from the library's static
scratch array and runs no tuning pass; the iteration kernels give the same
results on every tile, so the two entries must agree: equal status, iteration
and lambda-update counts, and the unknown within 1e-12 relative (the FP64
parity bar is 1e-10).  Two calls in a row, the second on a smaller grid, run
each case, in both orders with respect to the tuner (before and after the
grid's tile is cached), and a grid over the scratch limit is refused.
"""
import ctypes as C

import numpy as np
import pytest

from helpers import rel_err

pytestmark = pytest.mark.gpu


def _inputs(grid, transient, seed):
    rng = np.random.default_rng(seed)
    nx, ny, nz = grid
    f = lambda a: np.asfortranarray(a)
    X = f(rng.uniform(-1.0, 1.0, (nx + 2, ny + 2, nz + 2)))
    R = f(rng.uniform(-50.0, 50.0, (nx, ny, nz)))
    if not transient:
        R = f(R - R.mean())          # compatible source of the all-Neumann problem
    k = [f(10.0 ** rng.uniform(-0.5, 0.5, s)) for s in ((nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1))]
    return X, R, k


def _solve(ht, entry, grid, transient, host):
    from hydrothermal3d_amd import _lib
    X, R, k = host
    d = [ht.DeviceArray.from_host(a) for a in (X, R, *k)]
    h = (1.0 / grid[0], 0.5 / grid[1], 1.0 / grid[2])
    bct, bcv = ((0, 0, 1, 1, 0, 0), (0.0, 0.0, 1.0, 0.0, 0.0, 0.0)) if transient else ((0,) * 6, (0.0,) * 6)
    f = _lib.DyrelField(grid[0], grid[1], grid[2], h[0], h[1], h[2], d[0].ptr, d[1].ptr, d[2].ptr, d[3].ptr, d[4].ptr,
                        int(transient), float(transient), 2e-3, (C.c_int * 6)(*bct), (C.c_double * 6)(*bcv))
    p = ht.DyrelParams(nitmax=20000, nitout=100, ncheck=25, norm=1, tol=1e-10).c()
    rep = _lib.DyrelReport()
    st = getattr(_lib.lib(), entry)(C.byref(f), C.byref(p), C.byref(rep))
    return st, rep.as_dict(), d[0].get()


CASES = [((40, 24, 36), 1, 1), ((20, 12, 16), 0, 2)]   # the second call is on a smaller grid


@pytest.mark.parametrize("noalloc_first", [True, False], ids=["untuned", "tuned"])
def test_noalloc_matches_dyrel_solve(ht, noalloc_first):
    for grid, transient, seed in CASES:
        host = _inputs(grid, transient, seed + (0 if noalloc_first else 10))
        order = ["ht3d_dyrel_solve_noalloc", "ht3d_dyrel_solve"]
        if not noalloc_first:
            order.reverse()
        out = {e: _solve(ht, e, grid, transient, host) for e in order}
        (sa, ra, Xa), (sb, rb, Xb) = out["ht3d_dyrel_solve"], out["ht3d_dyrel_solve_noalloc"]
        assert sa == sb == 0, (grid, sa, sb)
        for key in ("iters", "n_checks", "n_lambda", "lmax"):
            assert ra[key] == rb[key], (grid, key, ra[key], rb[key])
        for key in ("lmin", "nF_abs"):
            assert abs(ra[key] - rb[key]) <= 1e-12 * abs(ra[key]), (grid, key)
        assert rel_err(Xb, Xa) <= 1e-12, (grid, rel_err(Xb, Xa))


def test_noalloc_refuses_grids_over_the_limit(ht):
    from hydrothermal3d_amd import _lib
    f = _lib.DyrelField(129, 128, 128, 0.1, 0.1, 0.1, None, None, None, None, None, 0, 0.0, 1.0,
                        (C.c_int * 6)(*(0,) * 6), (C.c_double * 6)(*(0.0,) * 6))
    p = ht.DyrelParams().c()
    rep = _lib.DyrelReport()
    assert _lib.lib().ht3d_dyrel_solve_noalloc(C.byref(f), C.byref(p), C.byref(rep)) == _lib.ERR_ARG
    assert b"scratch limit" in _lib.lib().ht3d_last_error()
