"""Second restatement of the dimensional (v18) property kernels, in NumPy.

TEST INFRASTRUCTURE ONLY (it verifies the C oracle), written from the
reference independently of ht3d_oracle_dim.c: whole-array slices with 0-based
offsets, and Dekker's exact product (Veltkamp splitting) where the C oracle
uses fma.  Same left-to-right evaluation order, so results equal the C oracle
bit for bit except where exp() is involved (NumPy has its own exp).

The ghost fills are restated differently on purpose: `SetTemperatureBCs_serial`
and `SetPressureBCs_serial` execute the reference kernel body literally, one
index at a time in ParallelStencil's loop order (x fastest), which is what the
reference's Threads backend does on one thread.  They are slow pure-Python
loops for small grids only.

Sources: Dimensional/HydroThermal3D_v18.jl:29-39 and
Dimensional/kernels_HT3D_dimensional.jl:40-119,148-160,183-205,225-259,291-314,328-332.
"""
from __future__ import annotations

import numpy as np

_SPLIT = 134217729.0   # 2^27 + 1


def _two_prod(a, b):
    """a*b = p + e exactly (Dekker 1971), without fma."""
    p = a * b
    c = _SPLIT * a
    ah = c - (c - a)
    al = a - ah
    c = _SPLIT * b
    bh = c - (c - b)
    bl = b - bh
    e = ((ah * bh - p) + ah * bl + al * bh) + al * bl
    return p, e


def pow4(x):
    """Julia's Float64 x^4: two squarings carrying their rounding errors."""
    x2, e2 = _two_prod(x, x)
    x4, e4 = _two_prod(x2, x2)
    return x4 + (e4 + (x2 * 2.0) * e2)


# ----------------------------------------------------- closures, v18:29-39
def kTs(T):
    return 3.1138 - 0.0023 * T


def kTf(T):
    return -0.069 + 0.0012 * T


def Cf(T):
    return 1000.0 + 7.5 * T


def Cs(T):
    return 0.5915 * T + 636.14


def rhos(T):
    return 2800.0 * (1.0 - (0.000024 * (T - 293.0)))


def rhof_C(Tc, p):
    terms = (7.424e-7 * p, -0.3922 * Tc, -4.441e-15 * (p * p), (4.547e-9 * p) * Tc, -0.003774 * (Tc * Tc),
             1.451e-23 * ((p * p) * p), (-1.793e-17 * (p * p)) * Tc, (7.485e-12 * p) * (Tc * Tc),
             2.955e-6 * ((Tc * Tc) * Tc), -1.463e-32 * pow4(p), (1.361e-26 * ((p * p) * p)) * Tc,
             (4.018e-21 * (p * p)) * (Tc * Tc), (-7.372e-15 * p) * ((Tc * Tc) * Tc), 5.698e-11 * pow4(Tc))
    r = 1006.0
    for t in terms:
        r = r + t
    return r


def drhodP_C(Tc, p):
    r = -7.372e-15 * ((Tc * Tc) * Tc)
    r = r + (8.036e-21 * (Tc * Tc)) * p
    r = r + 7.485e-12 * (Tc * Tc)
    r = r + (4.083e-26 * Tc) * (p * p)
    r = r - (3.586e-17 * Tc) * p
    r = r + 4.547e-9 * Tc
    r = r - 5.852e-32 * ((p * p) * p)
    r = r + 4.353e-23 * (p * p)
    r = r - 8.882e-15 * p
    return r + 7.424e-7


def muf(T):
    return 2.414e-5 * np.power(10.0, 247.8 / (T - 140.0))


def kF(y, delta):
    return 5e-16 * np.exp(y / delta)


def phi(phase, phi0):
    return np.where(phase == 2.0, 3.0 * phi0, phi0)


# ----------------------------------------------------------- kernels
def _vertex_mean(A):
    """Eight extended centroids around each vertex, the upper z plane first
    (kernels_HT3D_dimensional.jl:42-43)."""
    up = ((A[1:, 1:, 1:] + A[1:, :-1, 1:]) + A[:-1, 1:, 1:]) + A[:-1, :-1, 1:]
    lo = ((A[1:, 1:, :-1] + A[1:, :-1, :-1]) + A[:-1, 1:, :-1]) + A[:-1, :-1, :-1]
    return 0.125 * up + 0.125 * lo


def _cube_mean(A):
    """ParallelStencil @av: x, then y, then z neighbours, times 0.125."""
    return (A[:-1, :-1, :-1] + A[1:, :-1, :-1] + A[:-1, 1:, :-1] + A[1:, 1:, :-1] +
            A[:-1, :-1, 1:] + A[1:, :-1, 1:] + A[:-1, 1:, 1:] + A[1:, 1:, 1:]) * 0.125


def ComputeThermalConductivity(Tce, phv, phi0, scale_kt, scale_T):
    Tv = _vertex_mean(Tce) * scale_T
    f = phi(phv, phi0)
    return np.asfortranarray(((1.0 - f) * kTs(Tv) + f * kTf(Tv)) / scale_kt)


def ComputeHydroConductivity(kfv, Tce, Pce, phv, ymin, dy, k_fact, delta, s_sigma, s_T, s_L, s_t, s_eta, mode):
    Tv = _vertex_mean(Tce) * s_T
    Tv = np.where(Tv < 293.15, 293.15, Tv)
    Pv = _vertex_mean(Pce) * s_sigma
    y = (ymin + np.arange(Tv.shape[1], dtype=np.float64) * dy)[None, :, None]
    kf = kF(y * s_L, delta * s_L)
    out = np.array(kfv, dtype=np.float64, order="F", copy=True)
    if mode == 1:
        out[...] = rhof_C(Tv - 273.15, Pv) * kf / muf(Tv) / s_t
    elif mode == 2:
        out[...] = (kf / muf(Tv)) / (s_L * s_L / s_eta)
    elif mode == 3:
        out[...] = np.broadcast_to(kf / (s_L * s_L), out.shape)
    out[phv == 2.0] *= k_fact
    return out


def ComputeFluidDensity(Tce, Pce, phv, s_sigma, s_T, s_rho):
    Tv = _vertex_mean(Tce) * s_T
    Pv = _vertex_mean(Pce) * s_sigma
    out = rhof_C(Tv - 273.15, Pv) / s_rho
    out[phv == 1.0] = 0.0 / s_rho
    return np.asfortranarray(out)


def SmoothConductivityV2C(ktv):
    return np.asfortranarray(_cube_mean(ktv))


def SmoothConductivityC2V(ktv, ktc):
    out = np.array(ktv, order="F", copy=True)
    out[1:-1, 1:-1, 1:-1] = _cube_mean(ktc)
    return out


def ComputeRhoCeffective(rhoC, T, P, ph, phi_i, s_T, s_sigma, s_C, s_rho):
    I = slice(1, -1)
    p = ph[I, I, I]
    Ts, Ps = T[I, I, I] * s_T, P[I, I, I] * s_sigma
    f = phi(p, phi_i)
    rho = f * rhof_C(Ts - 273.15, Ps) + (1.0 - f) * rhos(Ts)
    C = f * Cf(Ts) + (1.0 - f) * Cs(Ts)
    out = np.array(rhoC, order="F", copy=True)
    m = np.abs(p) != 1.0
    out[m] = ((rho * C) / s_rho / s_C)[m]
    return out


def PhiRhof(out, T, P, ph, phi_i, s_rho, s_T, s_sigma, deriv=False):
    """ϕρf (T, P extended) or ϕρf_old (T, P cell arrays); ϕdρdP! with deriv."""
    I = slice(1, -1)
    if T.shape != out.shape:
        T, P = T[I, I, I], P[I, I, I]
    p = ph[I, I, I]
    Tc, Pp = T * s_T - 273.15, P * s_sigma
    val = phi(p, phi_i) * drhodP_C(Tc, Pp) / (s_rho / s_sigma) if deriv else phi(p, phi_i) * rhof_C(Tc, Pp) / s_rho
    res = np.array(out, order="F", copy=True)
    m = p != 1.0
    res[m] = val[m]
    return res


def dim_Init_vel(qx, qy, qz, rhofv):
    from oracle.ht3d_numpy import av_xya, av_xza, av_yza
    return qx / av_yza(rhofv), qy / av_xza(rhofv), qz / av_xya(rhofv)


# ------------------------------------------------------ ghost fills, one thread
def SetTemperatureBCs_serial(T, ky, wN, ph, qyS, _dy, TN):
    """kernels_HT3D_dimensional.jl:183-205 executed index by index (1-based
    names as in the reference, converted when indexing)."""
    T = np.array(T, order="F", copy=True)
    n1, n2, n3 = T.shape
    for iz in range(1, n3 + 1):
        for iy in range(1, n2 + 1):
            for ix in range(1, n1 + 1):
                a = lambda i, j, k: (i - 1, j - 1, k - 1)
                if ph[a(ix, iy, iz)] == 1.0:
                    T[a(ix, iy, iz)] = TN
                if ix == 1 and 1 < iy <= n2 - 1 and 1 < iz <= n3:
                    T[a(1, iy, iz)] = T[a(2, iy, iz)]
                if ix == n1 and 1 < iy <= n2 - 1 and 1 < iz <= n3:
                    T[a(n1, iy, iz)] = T[a(n1 - 1, iy, iz)]
                if iy == 1 and 1 < ix <= n1 - 1 and 1 < iz <= n3 - 1:
                    kS = ky[ix - 2, 0, iz - 2]
                    T[a(ix, 1, iz)] = qyS / _dy / kS + T[a(ix, 2, iz)]
                if ph[a(ix, iy, iz)] == -1.0:
                    w = wN[ix - 1, iz - 1]
                    T[a(ix, iy, iz)] = (TN - (1.0 - w) * T[a(ix, iy - 1, iz)]) / w
                if iz == 1 and 1 < ix <= n1 - 1 and 1 < iy <= n2 - 1:
                    T[a(ix, iy, 1)] = T[a(ix, iy, 2)]
                if iz == n3 and 1 < ix <= n1 - 1 and 1 < iy <= n2 - 1:
                    T[a(ix, iy, n3)] = T[a(ix, iy, n3 - 1)]
    return T


def SetPressureBCs_serial(P, wN, ph, PN):
    """kernels_HT3D_dimensional.jl:291-314 executed index by index."""
    P = np.array(P, order="F", copy=True)
    n1, n2, n3 = P.shape
    for iz in range(1, n3 + 1):
        for iy in range(1, n2 + 1):
            for ix in range(1, n1 + 1):
                a = lambda i, j, k: (i - 1, j - 1, k - 1)
                if ph[a(ix, iy, iz)] == 1.0:
                    P[a(ix, iy, iz)] = PN
                if ix == 1 and 1 < iy <= n2 - 1 and 1 < iz <= n3 - 1:
                    P[a(1, iy, iz)] = P[a(2, iy, iz)]
                if ix == n1 and 1 < iy <= n2 - 1 and 1 < iz <= n3 - 1:
                    P[a(n1, iy, iz)] = P[a(n1 - 1, iy, iz)]
                if iy == 1 and 1 < ix <= n1 - 1 and 1 < iz <= n3 - 1:
                    P[a(ix, 1, iz)] = P[a(ix, 2, iz)]
                if ph[a(ix, iy, iz)] == -1.0:
                    w = wN[ix - 1, iz - 1]
                    P[a(ix, iy, iz)] = (PN - (1.0 - w) * P[a(ix, iy - 1, iz)]) / w
                if iz == 1 and 1 < ix <= n1 - 1 and 1 < iy <= n2 - 1:
                    P[a(ix, iy, 1)] = P[a(ix, iy, 2)]
                if iz == n3 and 1 < ix <= n1 - 1 and 1 < iy <= n2 - 1:
                    P[a(ix, iy, n3)] = P[a(ix, iy, n3 - 1)]
    return P


# ------------------------------------------------------ v18 residuals and fluxes
def ComputeFlux(kx, ky, kz, A, _dx, _dy, _dz):
    """kernels_HT3D_dimensional.jl:140-146"""
    I = slice(1, -1)
    return (-kx * (_dx * (A[1:, I, I] - A[:-1, I, I])),
            -ky * (_dy * (A[I, 1:, I] - A[I, :-1, I])),
            -kz * (_dz * (A[I, I, 1:] - A[I, I, :-1])))


def ComputeDarcyFlux(rhofv, kx, ky, kz, P, g, _dx, _dy, _dz):
    """kernels_HT3D_dimensional.jl:266-272 (rho_f averaged to the y faces)"""
    from oracle.ht3d_numpy import av_xza
    I = slice(1, -1)
    return (-kx * (_dx * (P[1:, I, I] - P[:-1, I, I])),
            -ky * (_dy * (P[I, 1:, I] - P[I, :-1, I]) - g * av_xza(rhofv)),
            -kz * (_dz * (P[I, I, 1:] - P[I, I, :-1])))


def _flux_div(F, qx, qy, qz, _dx, _dy, _dz):
    F = F + (qx[1:] - qx[:-1]) * _dx
    F = F + (qy[:, 1:] - qy[:, :-1]) * _dy
    return F + (qz[:, :, 1:] - qz[:, :, :-1]) * _dz


def ResidualTemperatureLinearised(T, Told, rhoC, ph, PC, qx, qy, qz, Qt, transient, _dt, _dx, _dy, _dz):
    """kernels_HT3D_dimensional.jl:162-181"""
    I = slice(1, -1)
    p = ph[I, I, I]
    F = transient * rhoC * _dt * (T[I, I, I] - Told)
    F = _flux_div(F, qx, qy, qz, _dx, _dy, _dz)
    F = F - transient * np.where(p == 3.0, Qt, 0.0)
    F = F / PC
    return np.asfortranarray(np.where(np.abs(p) != 1.0, F, 0.0))


def ResidualFluidPressure(ph, phirhof, phirhof0, PC, qx, qy, qz, transient, _dt, _dx, _dy, _dz):
    """kernels_HT3D_dimensional.jl:274-289"""
    I = slice(1, -1)
    p = ph[I, I, I]
    F = transient * _dt * (phirhof - phirhof0)
    F = _flux_div(F, qx, qy, qz, _dx, _dy, _dz) / PC
    return np.asfortranarray(np.where(np.abs(p) != 1.0, F, 0.0))


# ------------------------------------------------------ output fields (kernels_HT3D_postprocessing.jl)
def _inner(A):
    return A[1:-1, 1:-1, 1:-1]


def _air(ph):
    return _inner(ph) == 1.0


def Phase(ph):
    return np.asfortranarray(np.where(_air(ph), np.nan, _inner(ph)))


def Scaled(F, ph, scale):
    """Pressure! / Temperature!"""
    return np.asfortranarray(np.where(_air(ph), np.nan, _inner(F) * scale))


def Velocity(X, Y, Z, Vx, Vy, Vz, ph, sV):
    X, Y, Z = (np.array(a, order="F", copy=True) for a in (X, Y, Z))
    air = _air(ph)
    vx = 0.5 * (Vx[:-1] + Vx[1:]) * sV
    vy = 0.5 * (Vy[:, :-1] + Vy[:, 1:]) * sV
    vz = 0.5 * (Vz[:, :, :-1] + Vz[:, :, 1:]) * sV
    X[...] = np.where(air, np.nan, vx)
    Y[~air], Z[~air] = vy[~air], vz[~air]
    return X, Y, Z


def EffectiveThermalConductivity(T, phi0, ph, sT):
    Ts = _inner(T) * sT
    f = phi(_inner(ph), phi0)
    return np.asfortranarray(np.where(_air(ph), np.nan, (1.0 - f) * kTs(Ts) + f * kTf(Ts)))


def EffectiveHeatCapacity(C, T, ph, sT):
    out = np.array(C, order="F", copy=True)
    m = ~_air(ph)
    out[m] = Cs(_inner(T) * sT)[m]
    return out


def Permeability(k_fact, ymin, dy, ph, delta, sL, shape):
    y = (ymin + np.arange(shape[1], dtype=np.float64) * dy + dy / 2.0)[None, :, None]
    v = np.broadcast_to(kF(y * sL, delta * sL), shape)
    v = np.where(_inner(ph) == 2.0, v * k_fact, v)
    return np.asfortranarray(np.where(_air(ph), np.nan, v))


def FluidDensity(T, P, ph, sT, ss):
    return np.asfortranarray(np.where(_air(ph), np.nan, rhof_C(_inner(T) * sT - 273.15, _inner(P) * ss)))


def Viscosity(T, ph, sT):
    return np.asfortranarray(np.where(_air(ph), np.nan, muf(_inner(T) * sT)))


def Peclet(T, P, Vx, Vy, Vz, ph, phi0, sT, ss, sV, sL):
    x = 0.5 * (Vx[:-1] + Vx[1:]) * sV
    y = 0.5 * (Vy[:, :-1] + Vy[:, 1:]) * sV
    z = 0.5 * (Vz[:, :, :-1] + Vz[:, :, 1:]) * sV
    V = np.sqrt(x * x + y * y + z * z)
    Ts, Ps = _inner(T) * sT, _inner(P) * ss
    f = phi(_inner(ph), phi0)
    keff = (1.0 - f) * kTs(Ts) + f * kTf(Ts)
    return np.asfortranarray(np.where(_air(ph), np.nan, sL * rhof_C(Ts - 273.15, Ps) * Cf(Ts) * V / keff))
