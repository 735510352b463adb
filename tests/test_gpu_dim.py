"""Dimensional (v18) property kernels and immersed-surface ghost fills on the
GPU against the CPU oracle on the same inputs: bit for bit, except the kernels
that call 10^x or exp() (the hydraulic conductivity), which agree to 1e-14
relative.  Grids leave partial blocks in every direction."""
import numpy as np
import pytest

from helpers import dim_state

pytestmark = pytest.mark.gpu

GRIDS = [(6, 5, 4), (70, 37, 19), (133, 20, 9)]


def _dev(ht, a):
    return ht.DeviceArray.from_host(np.asfortranarray(a))


@pytest.mark.parametrize("grid", GRIDS, ids=lambda g: "x".join(map(str, g)))
def test_property_kernels(ht, orc, grid):
    rng = np.random.default_rng(100 + grid[0])
    d = dim_state(*grid, rng, air_from=grid[1] - 1)
    sc, sh = d["sc"], d["sh"]
    K = ht.kernels
    Td, Pd, phv, phc = _dev(ht, d["T"]), _dev(ht, d["P"]), _dev(ht, d["phv"]), _dev(ht, d["phc"])

    print(f"ComputeThermalConductivity {grid}", flush=True)
    ktv = np.zeros(sh["V"], order="F")
    orc.ComputeThermalConductivity(ktv, d["T"], d["phv"], 0.01, sc.kt, sc.T)
    g = ht.DeviceArray.zeros(sh["V"])
    K.ComputeThermalConductivity(g, Td, phv, 0.01, sc.kt, sc.T)
    assert np.array_equal(g.get(), ktv)

    print("SmoothConductivity", flush=True)
    ktc = np.zeros(sh["c"], order="F")
    orc.SmoothConductivityV2C(ktc, ktv)
    gc = ht.DeviceArray.zeros(sh["c"])
    K.SmoothConductivityV2C(gc, g)
    assert np.array_equal(gc.get(), ktc)
    orc.SmoothConductivityC2V(ktv, ktc)
    K.SmoothConductivityC2V(g, gc)
    assert np.array_equal(g.get(), ktv)

    print("ComputeHydroConductivity", flush=True)
    for mode in (1, 2, 3):
        k0 = np.asfortranarray(rng.uniform(1.0, 2.0, sh["V"]))
        args = (-30e3 / sc.L, 35e3 / sc.L / grid[1], 50.0, 3000.0 / sc.L, sc.sigma, sc.T, sc.L, sc.t, sc.eta, mode)
        kc = k0.copy(order="F")
        orc.ComputeHydroConductivity(kc, d["T"], d["P"], d["phv"], *args)
        kg = _dev(ht, k0)
        K.ComputeHydroConductivity(kg, Td, Pd, phv, *args)
        np.testing.assert_allclose(kg.get(), kc, rtol=1e-14, atol=0)

    print("ComputeFluidDensity", flush=True)
    rf = np.zeros(sh["V"], order="F")
    orc.ComputeFluidDensity(rf, d["T"], d["P"], d["phv"], sc.sigma, sc.T, sc.rho)
    rg = ht.DeviceArray.zeros(sh["V"])
    K.ComputeFluidDensity(rg, Td, Pd, phv, sc.sigma, sc.T, sc.rho)
    assert np.array_equal(rg.get(), rf)

    print("ComputeRhoCeffective", flush=True)
    r0 = np.asfortranarray(rng.uniform(-1, 1, sh["c"]))
    rc = r0.copy(order="F")
    orc.ComputeRhoCeffective(rc, d["T"], d["P"], d["phc"], 0.01, sc.T, sc.sigma, sc.C, sc.rho)
    rgd = _dev(ht, r0)
    K.ComputeRhoCeffective(rgd, Td, Pd, phc, 0.01, sc.T, sc.sigma, sc.C, sc.rho)
    assert np.array_equal(rgd.get(), rc)

    print("PhiRhof / PhiRhof_old / PhiDrhoDP", flush=True)
    Tc0 = np.asfortranarray(d["T"][1:-1, 1:-1, 1:-1])
    Pc0 = np.asfortranarray(d["P"][1:-1, 1:-1, 1:-1])
    for fn, T, P in (("PhiRhof", d["T"], d["P"]), ("PhiRhof_old", Tc0, Pc0), ("PhiDrhoDP", d["T"], d["P"])):
        o0 = np.asfortranarray(rng.uniform(-1, 1, sh["c"]))
        oc = o0.copy(order="F")
        getattr(orc, fn)(oc, T, P, d["phc"], 0.01, sc.rho, sc.T, sc.sigma)
        og = _dev(ht, o0)
        getattr(K, fn)(og, _dev(ht, T), _dev(ht, P), phc, 0.01, sc.rho, sc.T, sc.sigma)
        assert np.array_equal(og.get(), oc), fn

    print("dim_Init_vel", flush=True)
    q = [np.asfortranarray(rng.uniform(-1, 1, sh[s])) for s in ("Fx", "Fy", "Fz")]
    V = [np.zeros(sh[s], order="F") for s in ("Fx", "Fy", "Fz")]
    orc.dim_Init_vel(*V, *q, rf + 1.0)
    Vg = [ht.DeviceArray.zeros(sh[s]) for s in ("Fx", "Fy", "Fz")]
    K.dim_Init_vel(*Vg, *[_dev(ht, a) for a in q], _dev(ht, rf + 1.0))
    for a, b in zip(Vg, V):
        assert np.array_equal(a.get(), b)


@pytest.mark.parametrize("immersed", [False, True], ids=["flat", "immersed"])
@pytest.mark.parametrize("grid", GRIDS, ids=lambda g: "x".join(map(str, g)))
def test_ghost_fills(ht, orc, grid, immersed):
    nx, ny, nz = grid
    rng = np.random.default_rng(200 + nx + immersed)
    d = dim_state(nx, ny, nz, rng, air_from=ny - 1, immersed=immersed)
    sh = d["sh"]
    ky = np.asfortranarray(rng.uniform(0.5, 2.0, sh["Fy"]))
    qyS, _dy = 37e-3 / d["sc"].W * d["sc"].L ** 2, ny / 3.5
    targs = (qyS, _dy, 1.05, 0.0, d["TN"]) + (0.0,) * 10
    print(f"dim_SetTemperatureBCs {grid} immersed={immersed}", flush=True)
    T = d["T"].copy(order="F")
    orc.dim_SetTemperatureBCs(T, ky, d["wN"], d["phc"], *targs)
    Tg = _dev(ht, d["T"])
    ht.kernels.dim_SetTemperatureBCs(Tg, _dev(ht, ky), _dev(ht, d["wN"]), _dev(ht, d["phc"]), *targs)
    assert np.array_equal(Tg.get(), T)
    print(f"dim_SetPressureBCs {grid} immersed={immersed}", flush=True)
    P = d["P"].copy(order="F")
    pargs = (0.0, d["PN"]) + (0.0,) * 8
    orc.dim_SetPressureBCs(P, d["wN"], d["phc"], *pargs)
    Pg = _dev(ht, d["P"])
    ht.kernels.dim_SetPressureBCs(Pg, _dev(ht, d["wN"]), _dev(ht, d["phc"]), *pargs)
    assert np.array_equal(Pg.get(), P)


def test_ghost_fill_rejects_bad_weight_table(ht):
    d = dim_state(6, 5, 4, np.random.default_rng(0), air_from=4)
    with pytest.raises(ValueError):
        ht.kernels.dim_SetPressureBCs(_dev(ht, d["P"]), _dev(ht, d["wN"][:, :-1]), _dev(ht, d["phc"]),
                                      *((0.0,) * 10))


@pytest.mark.parametrize("grid", GRIDS[:2], ids=lambda g: "x".join(map(str, g)))
def test_output_fields(ht, orc, grid):
    """kernels_HT3D_postprocessing.jl on the GPU against the oracle: bit for bit
    except Permeability and Viscosity (exp, 10^x: 1e-14 relative)."""
    rng = np.random.default_rng(300 + grid[0])
    d = dim_state(*grid, rng, air_from=grid[1] - 1)
    sc, sh = d["sc"], d["sh"]
    K = ht.kernels
    Td, Pd, phc = _dev(ht, d["T"]), _dev(ht, d["P"]), _dev(ht, d["phc"])
    V = [np.asfortranarray(rng.uniform(-1, 1, sh[s])) for s in ("Fx", "Fy", "Fz")]
    Vd = [_dev(ht, a) for a in V]
    X, Xd = np.zeros(sh["c"], order="F"), ht.DeviceArray.zeros(sh["c"])
    eq = lambda a, b: np.array_equal(a, b, equal_nan=True)
    print(f"output fields {grid}", flush=True)
    orc.Phase(X, d["phc"]); K.Phase(Xd, phc); assert eq(Xd.get(), X)
    orc.Pressure(X, d["P"], d["phc"], sc.sigma); K.Pressure(Xd, Pd, phc, sc.sigma); assert eq(Xd.get(), X)
    orc.Temperature(X, d["T"], d["phc"], sc.T); K.Temperature(Xd, Td, phc, sc.T); assert eq(Xd.get(), X)
    Y0 = [np.asfortranarray(rng.uniform(-1, 1, sh["c"])) for _ in range(3)]
    Yh = [a.copy(order="F") for a in Y0]
    Yd = [_dev(ht, a) for a in Y0]
    orc.Velocity(*Yh, *V, d["phc"], sc.V); K.Velocity(*Yd, *Vd, phc, sc.V)
    for a, b in zip(Yd, Yh):
        assert eq(a.get(), b)
    orc.EffectiveThermalConductivity(X, d["T"], 0.01, d["phc"], sc.T)
    K.EffectiveThermalConductivity(Xd, Td, 0.01, phc, sc.T); assert eq(Xd.get(), X)
    orc.EffectiveHeatCapacity(X, d["T"], d["P"], d["phc"], 0.01, sc.T, sc.sigma, sc.C, sc.rho)
    K.EffectiveHeatCapacity(Xd, Td, Pd, phc, 0.01, sc.T, sc.sigma, sc.C, sc.rho); assert eq(Xd.get(), X)
    orc.FluidDensity(X, d["T"], d["P"], d["phc"], sc.T, sc.sigma)
    K.FluidDensity(Xd, Td, Pd, phc, sc.T, sc.sigma); assert eq(Xd.get(), X)
    orc.Peclet(X, d["T"], d["P"], *V, d["phc"], 0.01, sc.T, sc.sigma, sc.V, sc.L)
    K.Peclet(Xd, Td, Pd, *Vd, phc, 0.01, sc.T, sc.sigma, sc.V, sc.L); assert eq(Xd.get(), X)
    orc.Permeability(X, 50.0, -30e3 / sc.L, 5.0, d["phc"], 3000.0 / sc.L, sc.L)
    K.Permeability(Xd, 50.0, -30e3 / sc.L, 5.0, phc, 3000.0 / sc.L, sc.L)
    np.testing.assert_allclose(Xd.get(), X, rtol=1e-14)
    orc.Viscosity(X, d["T"], d["phc"], sc.T)
    K.Viscosity(Xd, Td, phc, sc.T)
    np.testing.assert_allclose(Xd.get(), X, rtol=1e-14)
