"""Writes the golden fixtures of tests/golden from the NumPy restatement
(oracle/ht3d_numpy.py), which is independent of both the C oracle and the HIP
library.  Run from the repository root:  python tests/golden/make_golden.py

hrl_8x6x8.npz  closed-box HRL convection, 8x6x8 cells of the 2x1x2 box, Ra=500,
               lognormal k (sigma 0.5, 64 modes), 3 physical steps with DYREL
               (max-norm check every 25 iterations, lambda every 100, tol 1e-10)
hrl_visc_dtconst_8x6x8.npz
               the same with viscosity mu = exp(-ln(10) T) and a constant
               dt = 0.004 split into advection sub-steps
weno_7x5x6.npz one WENO5 advection step (v1 and v2) of a random field with random
               face velocities and the closed-box boundary types
"""
import math
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from hydrothermal3d_amd import inputs          # noqa: E402  (workload definitions and seeded inputs)
from oracle import ht3d_numpy as npo           # noqa: E402

HERE = Path(__file__).resolve().parent


def lognormal_numpy(w):
    """kfv = 10^(sigma sqrt(2/M) sum_m cos(K_m . x + phi_m)) at the vertices."""
    K, phi = inputs.lognormal_modes(w)
    dx, dy, dz = w.h
    x = np.arange(w.ncx + 1) * dx
    y = np.arange(w.ncy + 1) * dy
    z = np.arange(w.ncz + 1) * dz
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    s = np.zeros(X.shape)
    for m in range(w.modes):
        s += np.cos(K[m, 0] * X + K[m, 1] * Y + K[m, 2] * Z + phi[m])
    return np.asfortranarray(10.0 ** (w.sigma * (math.sqrt(2.0 / w.modes) * s)))


def hrl(name, mu_a=0.0, dt_constant=None):
    w = inputs.replace(inputs.WORKLOADS["hrl256"].scaled(8, 6, 8), modes=64)
    T0 = inputs.initial_temperature(w)
    kfv = lognormal_numpy(w)
    m = npo.HRL(w.ncx, w.ncy, w.ncz, w.Lx, w.Ly, w.Lz, w.Ra, w.cfl)
    m.T = T0.copy()
    m.kfv = kfv.copy()
    m.mu_a = mu_a
    if dt_constant is not None:
        m.dt_rule, m.dt = 1, dt_constant
    m.prm = dict(nitmax=20000, nitout=100, ncheck=25, norm=1, tol=1e-10)
    Nu, dt, itT, itP, nadv = [], [], [], [], []
    for _ in range(3):
        dt.append(m.dt)
        r = m.step()
        Nu.append(r["Nu"])
        itT.append(r["thermal"]["iters"])
        itP.append(r["darcy"]["iters"])
        nadv.append(r["n_adv"])
    np.savez_compressed(HERE / name, T0=T0, kfv=kfv, T=m.T, P=m.P, Vx=m.V[0], Vy=m.V[1], Vz=m.V[2],
                        Nu=np.array(Nu), dt=np.array(dt), iters_T=np.array(itT), iters_P=np.array(itP),
                        n_adv=np.array(nadv), dims=np.array([w.ncx, w.ncy, w.ncz]), L=np.array([w.Lx, w.Ly, w.Lz]),
                        Ra=w.Ra, cfl=w.cfl, modes=w.modes, seed=w.seed, mu_a=mu_a,
                        dt_constant=-1.0 if dt_constant is None else dt_constant)
    return nadv


def weno():
    rng = np.random.default_rng(2024)
    nx, ny, nz = 7, 5, 6
    T = np.asfortranarray(rng.uniform(0, 1, (nx + 2, ny + 2, nz + 2)))
    Vx = np.asfortranarray(rng.uniform(-1, 1, (nx + 1, ny, nz)))
    Vy = np.asfortranarray(rng.uniform(-1, 1, (nx, ny + 1, nz)))
    Vz = np.asfortranarray(rng.uniform(-1, 1, (nx, ny, nz + 1)))
    Vx.ravel(order="F")[::3] = 0.0
    args = (2.5e-2, 7.0, 5.0, 6.0, 0.0, 1.0)
    T1 = npo.advect_weno5(T, Vx, Vy, Vz, *args, version=1)
    T2 = npo.advect_weno5(T, Vx, Vy, Vz, *args, version=2)
    np.savez_compressed(HERE / "weno_7x5x6.npz", T=T, Vx=Vx, Vy=Vy, Vz=Vz, args=np.array(args), T_v1=T1, T_v2=T2)


if __name__ == "__main__":
    hrl("hrl_8x6x8.npz")
    print("sub-steps", hrl("hrl_visc_dtconst_8x6x8.npz", mu_a=math.log(10.0), dt_constant=0.004))
    weno()
    print("golden fixtures written to", HERE)
