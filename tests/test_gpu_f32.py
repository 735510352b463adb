"""FP32 inner-sweep kernels (the *_f32 entries) against the NumPy restatement
evaluated in float32 (IEEE single, one rounding per operation, the reference's
order), bit for bit, and the FP64/FP32 glue (to_f32, add_f32)."""
import numpy as np
import pytest

from helpers import shapes

pytestmark = pytest.mark.gpu

GRIDS = [(3, 3, 3), (13, 9, 11), (67, 6, 5)]
f32 = np.float32


def _r(shape, rng, lo=-1.0, hi=1.0):
    return np.asfortranarray(rng.uniform(lo, hi, shape).astype(np.float32))


def _dev(ht, a):
    return ht.DeviceArray.from_host(a)


@pytest.mark.parametrize("grid", GRIDS, ids=lambda g: "x".join(map(str, g)))
def test_f32_kernels_bitwise(ht, grid):
    from oracle import ht3d_numpy as npo
    rng = np.random.default_rng(5)
    sh = shapes(*grid)
    k = [_r(sh[s], rng, 0.1, 3.0) for s in ("Fx", "Fy", "Fz")]
    A, T, R, F0, PC = _r(sh["E"], rng), _r(sh["E"], rng), _r(sh["c"], rng, -5, 5), _r(sh["c"], rng), \
        _r(sh["c"], rng, 1.0, 4.0)
    rhoC = _r(sh["c"], rng, 0.5, 2.0)
    _dx, _dy, _dz, _dt = f32(17.0), f32(33.3), f32(9.1), f32(41.7)
    kd = [_dev(ht, a) for a in k]
    K = ht.kernels
    # ComputeFlux
    q = [ht.DeviceArray.zeros(sh[s], np.float32) for s in ("Fx", "Fy", "Fz")]
    K.ComputeFlux_f32(*q, *kd, _dev(ht, A), _dx, _dy, _dz)
    qn = npo.ComputeFlux(*k, A, _dx, _dy, _dz)
    for a, b in zip(q, qn):
        assert a.get().dtype == np.float32 and np.array_equal(a.get(), b)
    qd = [_dev(ht, np.asfortranarray(b)) for b in qn]
    # residuals, with and without the diagonal scaling
    F = ht.DeviceArray.zeros(sh["c"], np.float32)
    K.ResidualDiffusion_f32(F, _dev(ht, R), *qd, _dx, _dy, _dz)
    assert np.array_equal(F.get(), npo.ResidualDiffusion(R, *qn, _dx, _dy, _dz))
    K.ResidualDiffusionTransient_f32(F, _dev(ht, T), _dev(ht, R), *qd, _dt, _dx, _dy, _dz)
    Ft = npo.ResidualDiffusionTransient(T, R, *qn, _dt, _dx, _dy, _dz)
    assert np.array_equal(F.get(), Ft)
    K.ResidualDiffusionPC_f32(F, _dev(ht, R), *qd, _dev(ht, PC), _dx, _dy, _dz)
    assert np.array_equal(F.get(), npo.ResidualDiffusion(R, *qn, _dx, _dy, _dz) / PC)
    K.ResidualDiffusionTransientPC_f32(F, _dev(ht, T), _dev(ht, R), *qd, _dev(ht, PC), _dt, _dx, _dy, _dz)
    assert np.array_equal(F.get(), Ft / PC)
    # Gershgorin
    G, D = ht.DeviceArray.zeros(sh["c"], np.float32), ht.DeviceArray.zeros(sh["c"], np.float32)
    dx, dy, dz, dt = f32(1 / 17.0), f32(1 / 33.3), f32(1 / 9.1), f32(1 / 41.7)
    K.GershgorinPoisson_f32(G, D, _dev(ht, rhoC), *kd, f32(1.0), dt, dx, dy, dz)
    Gn, Dn = npo.GershgorinPoisson(rhoC, *k, f32(1.0), dt, dx, dy, dz)
    assert np.array_equal(D.get(), Dn) and np.array_equal(G.get(), Gn)
    # DYREL update
    Fd, Xd = _dev(ht, F0), _dev(ht, A)
    K.DYRELUpdate_f32(Fd, Xd, _dev(ht, Ft), f32(0.93), f32(1.7e-3), f32(2.1e-2))
    F0n, Xn = npo.DYRELUpdate(F0, A, Ft, f32(0.93), f32(1.7e-3), f32(2.1e-2))
    assert np.array_equal(Fd.get(), F0n) and np.array_equal(Xd.get(), Xn)
    # ghost fills
    Td = _dev(ht, T)
    K.SetTemperatureBCs_f32(Td, 1.0, 0.0)
    assert np.array_equal(Td.get(), npo.temperature_bc(T, f32(1.0), f32(0.0)))
    Pd = _dev(ht, A)
    K.SetPressureBCs_f32(Pd, 0.0, 0.0)
    assert np.array_equal(Pd.get(), npo.pressure_bc(A))


def test_f32_glue(ht):
    rng = np.random.default_rng(9)
    sh = shapes(10, 7, 5)
    X = np.asfortranarray(rng.standard_normal(sh["E"]))
    e = _r(sh["E"], rng)
    out = ht.DeviceArray.zeros(sh["E"], np.float32)
    ht.kernels.to_f32(out, _dev(ht, X))
    assert np.array_equal(out.get(), X.astype(np.float32))
    Xd = _dev(ht, X)
    ht.kernels.add_f32(Xd, _dev(ht, e))
    ref = X.copy()
    ref[1:-1, 1:-1, 1:-1] += e[1:-1, 1:-1, 1:-1].astype(np.float64)
    assert np.array_equal(Xd.get(), ref)
