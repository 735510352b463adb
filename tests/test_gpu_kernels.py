"""Per-kernel parity: every reference kernel of the HIP drop-in layer against the
CPU oracle on the same seeded inputs, bit for bit (FP64, no FMA contraction on
either side).  Grids are small and ragged (odd, unequal extents, down to the
minimum of 3 cells the reference's WENO halo needs)."""
import zlib

import numpy as np
import pytest

from helpers import shapes, rand, positive

pytestmark = pytest.mark.gpu

GRIDS = [(3, 3, 3), (7, 5, 6), (13, 9, 11), (67, 6, 5)]

# name -> (args spec, outputs); spec entries: (argname, kind) for arrays, or
# (argname, float/int) for scalars.  kind suffix "+" draws a positive array.
S = {
    "ResetA": ([("A", "c"), ("B", "c")], ["A", "B"]),
    "Cpy_inn_to_all": ([("A", "c"), ("B", "E")], ["A"]),
    "Cpy_all_to_inn": ([("A", "E"), ("B", "c")], ["A"]),
    "ArrayEqualArray": ([("A", "c"), ("B", "c")], ["A"]),
    "Multiply": ([("a", "c"), ("b", "c"), ("c", "c")], ["a"]),
    "InitThermal": ([("Rt", "c"), ("kx", "Fx"), ("ky", "Fy"), ("kz", "Fz"), ("Tc_ex", "E"), ("ktv", "V+"), ("_dt", 37.5)],
                    ["Rt", "kx", "ky", "kz"]),
    "InitConductivity": ([("kx", "Fx"), ("ky", "Fy"), ("kz", "Fz"), ("ktv", "V+")], ["kx", "ky", "kz"]),
    "InitDarcy": ([("Ty", "Fy"), ("kx", "Fx"), ("ky", "Fy"), ("kz", "Fz"), ("Tc_ex", "E"), ("kfv", "V+"), ("_dt", 3.0)],
                  ["Ty", "kx", "ky", "kz"]),
    "KillNorthSouthFluxes": ([("Ty", "Fy")], ["Ty"]),
    "Compute_Rp": ([("Rp", "c"), ("ky", "Fy+"), ("Ty", "Fy"), ("Ra", 517.25), ("_dy", 31.0)], ["Rp"]),
    "SetTemperatureBCs": ([("Tc_ex", "E"), ("Tbot", 1.0), ("Ttop", 0.0)], ["Tc_ex"]),
    "SetPressureBCs": ([("Pc_ex", "E"), ("Pbot", 0.25), ("Ptop", 0.0)], ["Pc_ex"]),
    "SetPressureBCs_v7": ([("Pc_ex", "E"), ("Pbot", 0.25), ("Ptop", -0.5)], ["Pc_ex"]),
    "ComputeFlux": ([("qx", "Fx"), ("qy", "Fy"), ("qz", "Fz"), ("kx", "Fx+"), ("ky", "Fy+"), ("kz", "Fz+"), ("A", "E"),
                     ("_dx", 17.0), ("_dy", 33.0), ("_dz", 9.0)], ["qx", "qy", "qz"]),
    "UpdateT": ([("F", "c"), ("T", "E"), ("R", "c"), ("qx", "Fx"), ("qy", "Fy"), ("qz", "Fz"), ("_dt", 41.0),
                 ("dtau", 1.3e-3), ("_dx", 17.0), ("_dy", 33.0), ("_dz", 9.0)], ["F", "T"]),
    "ResidualDiffusion": ([("F", "c"), ("R", "c"), ("qx", "Fx"), ("qy", "Fy"), ("qz", "Fz"), ("_dx", 17.0), ("_dy", 33.0),
                           ("_dz", 9.0)], ["F"]),
    "ResidualDiffusionTransient": ([("F", "c"), ("T", "E"), ("R", "c"), ("qx", "Fx"), ("qy", "Fy"), ("qz", "Fz"),
                                    ("_dt", 41.0), ("_dx", 17.0), ("_dy", 33.0), ("_dz", 9.0)], ["F"]),
    "ResidualDiffusionPC": ([("F", "c"), ("R", "c"), ("qx", "Fx"), ("qy", "Fy"), ("qz", "Fz"), ("PC", "c+"), ("_dx", 17.0),
                             ("_dy", 33.0), ("_dz", 9.0)], ["F"]),
    "ResidualDiffusionTransientPC": ([("F", "c"), ("T", "E"), ("R", "c"), ("qx", "Fx"), ("qy", "Fy"), ("qz", "Fz"),
                                      ("PC", "c+"), ("_dt", 41.0), ("_dx", 17.0), ("_dy", 33.0), ("_dz", 9.0)], ["F"]),
    "UpdateP": ([("F0", "c"), ("P", "E"), ("F", "c"), ("dampx", 0.93), ("dtau", 2.1e-4)], ["F0", "P"]),
    "DYRELUpdate": ([("F0", "c"), ("X", "E"), ("F", "c"), ("h1", 0.987), ("h2", 3.3e-3), ("dtau", 2.1e-3)], ["F0", "X"]),
    "SwapDYREL": ([("Xc", "c"), ("Yc", "c"), ("X0c_ex", "E"), ("Y0c", "c")], ["Xc", "Yc"]),
    "GershgorinPoisson": ([("G", "c"), ("D", "c"), ("rhoC", "c+"), ("kx", "Fx+"), ("ky", "Fy+"), ("kz", "Fz+"),
                           ("transient", 1.0), ("dt", 3.1e-3), ("dx", 0.1), ("dy", 0.07), ("dz", 0.13)], ["G", "D"]),
    "GershgorinHydro": ([("G", "c"), ("phidrhodP", "c+"), ("kx", "Fx+"), ("ky", "Fy+"), ("kz", "Fz+"), ("transient", 1.0),
                         ("dt", 3.1e-3), ("dx", 0.1), ("dy", 0.07), ("dz", 0.13)], ["G"]),
    "RayleighQuotientNumerator": ([("G", "c"), ("Fc0", "c"), ("Fcit", "c"), ("Fc", "c"), ("dtau", 7.7e-3)], ["G"]),
    "RayleighQuotientDenominator": ([("G", "c"), ("Fc0", "c"), ("dtau", 7.7e-3)], ["G"]),
    "ComputeDarcyFlux": ([("qx", "Fx"), ("qy", "Fy"), ("qz", "Fz"), ("rhofv", "V+"), ("kx", "Fx+"), ("ky", "Fy+"),
                          ("kz", "Fz+"), ("Pf", "E"), ("g", -9.81), ("_dx", 17.0), ("_dy", 33.0), ("_dz", 9.0)],
                         ["qx", "qy", "qz"]),
    "Init_vel": ([("Vx", "Fx"), ("Vy", "Fy"), ("Vz", "Fz"), ("kx", "Fx+"), ("ky", "Fy+"), ("kz", "Fz+"), ("Pc_ex", "E"),
                  ("Ty", "Fy"), ("Ra", 517.25), ("_dx", 17.0), ("_dy", 33.0), ("_dz", 9.0)], ["Vx", "Vy", "Vz"]),
    "VxPlusMinus": ([("Vxm", "c"), ("Vxp", "c"), ("Vx", "Fx")], ["Vxm", "Vxp"]),
    "VyPlusMinus": ([("Vym", "c"), ("Vyp", "c"), ("Vy", "Fy")], ["Vym", "Vyp"]),
    "VzPlusMinus": ([("Vzm", "c"), ("Vzp", "c"), ("Vz", "Fz")], ["Vzm", "Vzp"]),
    "dFdx_Weno5": ([("dFdxi", "c"), ("V1", "c"), ("V2", "c"), ("V3", "c"), ("V4", "c"), ("V5", "c")], ["dFdxi"]),
    "Advect": ([("Fc", "c"), ("Vxp", "c"), ("dTdxm", "c"), ("Vxm", "c"), ("dTdxp", "c"), ("dt", 1.7e-3)], ["Fc"]),
    "TimeAveraging": ([("Fc", "c"), ("Fold", "c"), ("order", 2.0)], ["Fc"]),
}


def _inputs(spec, grid, rng):
    sh = shapes(*grid)
    host = {}
    for name, kind in spec:
        if isinstance(kind, str):
            pos = kind.endswith("+")
            k = kind.rstrip("+")
            host[name] = positive(sh[k], rng) if pos else rand(sh[k], rng)
        else:
            host[name] = kind
    return host


def _run_pair(ht, orc, kname, spec, host):
    dev = {n: (ht.DeviceArray.from_host(v) if isinstance(v, np.ndarray) else v) for n, v in host.items()}
    cpu = {n: (v.copy(order="F") if isinstance(v, np.ndarray) else v) for n, v in host.items()}
    getattr(ht.kernels, kname)(*[dev[n] for n, _ in spec])
    getattr(orc, kname)(*[cpu[n] for n, _ in spec])
    return dev, cpu


@pytest.mark.parametrize("grid", GRIDS, ids=lambda g: "x".join(map(str, g)))
@pytest.mark.parametrize("kname", sorted(S))
def test_kernel_bitwise(ht, orc, kname, grid):
    spec, outs = S[kname]
    rng = np.random.default_rng([zlib.crc32(kname.encode()), *grid])
    host = _inputs(spec, grid, rng)
    if kname.endswith("PlusMinus"):
        # exercise the untouched branch (V == 0) and both signs
        vname = spec[2][0]
        host[vname].ravel(order="F")[::5] = 0.0
    dev, cpu = _run_pair(ht, orc, kname, spec, host)
    for o in outs:
        got, want = dev[o].get(), cpu[o]
        assert np.array_equal(got, want), f"{kname}.{o} differs: max |diff| {np.max(np.abs(got - want))}"


@pytest.mark.parametrize("grid", GRIDS[:3], ids=lambda g: "x".join(map(str, g)))
@pytest.mark.parametrize("axis", [1, 2, 3])
@pytest.mark.parametrize("plus", [0, 1])
def test_gradients_bitwise(ht, orc, grid, axis, plus):
    rng = np.random.default_rng(100 * axis + plus)
    sh = shapes(*grid)
    X6 = rand(sh["X6"], rng)
    vs_d = [ht.DeviceArray.zeros(sh["c"]) for _ in range(5)]
    vs_h = [np.zeros(sh["c"], order="F") for _ in range(5)]
    names = {(1, 0): "Gradients_minus_x_Weno5", (1, 1): "Gradients_plus_x_Weno5",
             (2, 0): "Gradients_minus_y_Weno5", (2, 1): "Gradients_plus_y_Weno5",
             (3, 0): "Gradients_minus_z_Weno5", (3, 1): "Gradients_plus_z_Weno5"}
    getattr(ht.kernels, names[(axis, plus)])(*vs_d, ht.DeviceArray.from_host(X6), 17.0, 33.0, 9.0)
    orc.Gradients_Weno5(*vs_h, X6, axis, plus, 17.0, 33.0, 9.0)
    for d, h in zip(vs_d, vs_h):
        assert np.array_equal(d.get(), h)


@pytest.mark.parametrize("grid", GRIDS[:3], ids=lambda g: "x".join(map(str, g)))
@pytest.mark.parametrize("axis", [1, 2, 3])
@pytest.mark.parametrize("types", [(0, 0), (1, 1), (2, 2), (0, 1)])
@pytest.mark.parametrize("src_ext", [0, 1])
def test_boundaries_bitwise(ht, orc, grid, axis, types, src_ext):
    rng = np.random.default_rng(7 * axis + 3 * types[0] + types[1] + 50 * src_ext)
    sh = shapes(*grid)
    X6 = rand(sh["X6"], rng)
    Fc = rand(sh["E"] if src_ext else sh["c"], rng)
    d = ht.DeviceArray.from_host(X6)
    fn = {1: ht.kernels.Boundaries_x_Weno5, 2: ht.kernels.Boundaries_y_Weno5, 3: ht.kernels.Boundaries_z_Weno5}[axis]
    fn(d, ht.DeviceArray.from_host(Fc), types[0], 1.25, types[1], -0.5)
    h = X6.copy(order="F")
    orc.Boundaries_Weno5(h, Fc, axis, types[0], 1.25, types[1], -0.5)
    assert np.array_equal(d.get(), h)


@pytest.mark.parametrize("grid", GRIDS[:3], ids=lambda g: "x".join(map(str, g)))
@pytest.mark.parametrize("axis", [1, 2, 3])
def test_advect_v2_bitwise(ht, orc, grid, axis):
    rng = np.random.default_rng(axis)
    sh = shapes(*grid)
    Fc = rand(sh["E"], rng)
    X6 = rand(sh["X6"], rng)
    V = rand(sh[{1: "Fx", 2: "Fy", 3: "Fz"}[axis]], rng)
    V.ravel(order="F")[::7] = 0.0
    d = ht.DeviceArray.from_host(Fc)
    ht.kernels.Advect(d, ht.DeviceArray.from_host(X6), ht.DeviceArray.from_host(V), axis, 1.3e-3, 17.0, 33.0, 9.0)
    h = Fc.copy(order="F")
    orc.Advect_v2(h, X6, V, axis, 1.3e-3, 17.0, 33.0, 9.0)
    assert np.array_equal(d.get(), h)


@pytest.mark.parametrize("grid", GRIDS[:3], ids=lambda g: "x".join(map(str, g)))
def test_timeaveraging_v2_bitwise(ht, orc, grid):
    rng = np.random.default_rng(5)
    sh = shapes(*grid)
    Fc, Fold = rand(sh["E"], rng), rand(sh["c"], rng)
    d = ht.DeviceArray.from_host(Fc)
    ht.kernels.TimeAveraging(d, ht.DeviceArray.from_host(Fold), 2.0)
    h = Fc.copy(order="F")
    orc.TimeAveraging_v2(h, Fold, 2.0)
    assert np.array_equal(d.get(), h)


@pytest.mark.parametrize("grid", GRIDS[:3], ids=lambda g: "x".join(map(str, g)))
def test_dimensional_residuals_bitwise(ht, orc, grid):
    """ResidualTemperatureLinearised! / ResidualFluidPressure! with air (|ph|==1),
    heat-source (ph==3) and rock phases mixed."""
    rng = np.random.default_rng(11)
    sh = shapes(*grid)
    ph = np.asfortranarray(rng.choice([-1.0, 1.0, 2.0, 3.0], size=sh["E"]))
    args = dict(T=rand(sh["E"], rng), T0=rand(sh["c"], rng), rc=positive(sh["c"], rng), PC=positive(sh["c"], rng),
                qx=rand(sh["Fx"], rng), qy=rand(sh["Fy"], rng), qz=rand(sh["Fz"], rng))
    D = {k: ht.DeviceArray.from_host(v) for k, v in args.items()}
    Fd, Fh = ht.DeviceArray.zeros(sh["c"]), np.zeros(sh["c"], order="F")
    phd = ht.DeviceArray.from_host(ph)
    ht.kernels.ResidualTemperatureLinearised(Fd, D["T"], D["T0"], D["rc"], phd, D["PC"], D["qx"], D["qy"], D["qz"],
                                             2.5, 1.0, 41.0, 17.0, 33.0, 9.0)
    orc.ResidualTemperatureLinearised(Fh, args["T"], args["T0"], args["rc"], ph, args["PC"], args["qx"], args["qy"],
                                      args["qz"], 2.5, 1.0, 41.0, 17.0, 33.0, 9.0)
    assert np.array_equal(Fd.get(), Fh)
    ht.kernels.ResidualFluidPressure(Fd, phd, D["rc"], D["T0"], D["PC"], D["qx"], D["qy"], D["qz"], 1.0, 41.0, 17.0,
                                     33.0, 9.0)
    orc.ResidualFluidPressure(Fh, ph, args["rc"], args["T0"], args["PC"], args["qx"], args["qy"], args["qz"], 1.0, 41.0,
                              17.0, 33.0, 9.0)
    assert np.array_equal(Fd.get(), Fh)


@pytest.mark.parametrize("grid", GRIDS[:3], ids=lambda g: "x".join(map(str, g)))
def test_viscosity_mobility(ht, orc, grid):
    """Face mobility k exp(a T_face): device exp against libm exp, so ulp-level."""
    rng = np.random.default_rng(31)
    sh = shapes(*grid)
    T = rand(sh["E"], rng, 0.0, 1.0)
    k = [positive(sh[s], rng) for s in ("Fx", "Fy", "Fz")]
    d = [ht.DeviceArray.from_host(a) for a in k]
    ht.kernels.ApplyViscosity(*d, ht.DeviceArray.from_host(T), 2.302585092994046)
    h = [a.copy(order="F") for a in k]
    orc.ApplyViscosity(*h, T, 2.302585092994046)
    for a, b in zip(d, h):
        assert np.allclose(a.get(), b, rtol=6e-16, atol=0)


@pytest.mark.parametrize("op", ["max", "min", "sum", "sum_abs", "max_abs"])
@pytest.mark.parametrize("n", [1, 63, 64, 1000, 300_001])
def test_reductions(ht, orc, op, n):
    rng = np.random.default_rng(n)
    a = rand((n, 1, 1), rng)
    d = ht.DeviceArray.from_host(a)
    code = {"sum": 0, "sum_abs": 1, "max": 2, "min": 3, "max_abs": 4}[op]
    got = ht.device.reduce(d, code)
    want = {"sum": a.sum(), "sum_abs": np.abs(a).sum(), "max": a.max(), "min": a.min(), "max_abs": np.abs(a).max()}[op]
    if op in ("max", "min", "max_abs"):
        assert got == want
    else:
        assert abs(got - want) <= 1e-13 * max(1.0, np.abs(a).sum())


@pytest.mark.parametrize("op, want", [(0, 0.0), (1, 0.0), (2, -np.inf), (3, np.inf), (4, -np.inf)],
                         ids=["sum", "sum_abs", "max", "min", "max_abs"])
def test_empty_reduction_is_the_identity(ht, op, want):
    """An empty array reduces to the identity of the operation, with no launch."""
    d = ht.DeviceArray.zeros((0,))
    assert ht.device.reduce(d, op) == want


@pytest.mark.parametrize("grid", [(2, 3, 3), (3, 2, 3), (3, 3, 2), (0, 5, 5)], ids=lambda g: "x".join(map(str, g)))
def test_grid_below_three_cells_is_refused(ht, grid):
    """Every kernel entry needs at least 3 cells per direction (the 3-layer WENO
    mirror and the ghost rules): a smaller grid is refused by the C-ABI with
    HT3D_ERR_ARG before any launch, and the library stays usable."""
    sh = shapes(*grid)
    q = [ht.DeviceArray.zeros(sh[k]) for k in ("Fx", "Fy", "Fz")]
    k = [ht.DeviceArray.zeros(sh[k]) for k in ("Fx", "Fy", "Fz")]
    A = ht.DeviceArray.zeros(sh["E"])
    with pytest.raises(ht._lib.HT3DError) as err:
        ht.kernels.ComputeFlux(*q, *k, A, 1.0, 1.0, 1.0)
    assert err.value.code == 1
    ok = shapes(3, 3, 3)
    ht.kernels.ComputeFlux(*[ht.DeviceArray.zeros(ok[s]) for s in ("Fx", "Fy", "Fz", "Fx", "Fy", "Fz")],
                           ht.DeviceArray.zeros(ok["E"]), 1.0, 1.0, 1.0)


def test_shape_mismatch_raises(ht):
    sh = shapes(5, 6, 7)
    bad = ht.DeviceArray.zeros((7, 6, 7))
    good = [ht.DeviceArray.zeros(sh[k]) for k in ("Fx", "Fy", "Fz", "Fx", "Fy", "Fz")]
    A = ht.DeviceArray.zeros(sh["E"])
    with pytest.raises(ValueError):
        ht.kernels.ComputeFlux(bad, *good[1:], A, 1.0, 1.0, 1.0)


@pytest.mark.parametrize("grid", GRIDS, ids=lambda g: "x".join(map(str, g)))
def test_velocity_dt_nu(ht, orc, grid):
    """Fused velocity + time step + Nu: V bit-identical to Init_vel (the oracle's
    kernels_HT3D.jl:73-80 restatement), max|V| exact, dt by the v18 rule, Nu as
    the oracle's top-face flux sum (HydroThermal3D_Alban_v6.jl:345-346)."""
    rng = np.random.default_rng(77)
    sh = shapes(*grid)
    nx, ny, nz = grid
    P, T = rand(sh["E"], rng), rand(sh["E"], rng, 0.0, 1.0)
    k = [positive(sh[s], rng) for s in ("Fx", "Fy", "Fz")]
    Ty, kty = rand(sh["Fy"], rng), positive(sh["Fy"], rng)
    Ra, h, cfl, Lx, Lz = 517.25, (0.7 / nx, 0.45 / ny, 1.3 / nz), 0.5, 0.7, 1.3
    dev = {n: ht.DeviceArray.from_host(a) for n, a in dict(P=P, T=T, kx=k[0], ky=k[1], kz=k[2], Ty=Ty, kty=kty).items()}
    V = [ht.DeviceArray.zeros(sh[s]) for s in ("Fx", "Fy", "Fz")]
    vmax, dt, nu = ht.kernels.velocity_dt_nu(*V, dev["kx"], dev["ky"], dev["kz"], dev["P"], dev["Ty"], Ra, *h, cfl,
                                             dev["T"], dev["kty"], Lx, Lz)
    Vo = [np.zeros(sh[s], order="F") for s in ("Fx", "Fy", "Fz")]
    orc.Init_vel(*Vo, *k, P, Ty, Ra, 1.0 / h[0], 1.0 / h[1], 1.0 / h[2])
    for a, b in zip(V, Vo):
        assert np.array_equal(a.get(), b)
    assert vmax == max(np.abs(v).max() for v in Vo)
    assert dt == cfl * min(h) / vmax
    flux = -kty[:, ny, :] * ((1.0 / h[1]) * (T[1:-1, ny + 1, 1:-1] - T[1:-1, ny, 1:-1]))
    import math
    assert abs(nu - math.fsum(flux.ravel()) / (Lx * Lz) / (nx * nz)) <= 1e-15 * abs(nu)
