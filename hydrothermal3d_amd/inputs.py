"""Workloads and seeded synthetic inputs for closed-box porous convection.

Axes follow the reference: x, y (vertical: gravity, Dirichlet T, Ty, Rp) and z,
x fastest (Dimensionless/kernels_HT3D.jl:76,94).  A workload quoted as
"a x b x c cells, domain 2x2x1" with a vertical third axis maps to reference
extents (a, c, b) and lengths (2, 1, 2).

Random numbers come from NumPy's counter-based Philox-4x32-10 bit generator
keyed by the workload seed, so every run of a workload sees the same inputs.

* Initial temperature: T = (1 - y/Ly) + 0.01 (2u - 1) at the interior cell
  centres, u ~ U[0,1) in x-fastest cell order; ghost cells carry the linear
  profile (HydroThermal3D_Alban_v6.jl:169 sets the whole Tc_ex from y).
* Permeability (vertex array kfv): 1, or log10 k = sigma G(x) with G a
  unit-variance Gaussian field of exponential covariance exp(-r/l), built by
  spectral randomisation: G(x) = sqrt(2/M) sum_m cos(K_m . x + phi_m) with
  phi ~ U[0, 2pi) and K_m = Z_m / (l |g_m|), Z ~ N(0, I3), g ~ N(0, 1) -- the
  multivariate Cauchy law, which is the normalised spectral density
  (1 + l^2 |K|^2)^-2 of exp(-r/l) in 3D.  Optional planar fault: |n.(x-x0)| <=
  2.5 h with n = (sin dip, cos dip, 0) through the domain centre, k x 100.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, replace

import numpy as np

from . import _lib
from .device import DeviceArray


@dataclass(frozen=True)
class Workload:
    name: str
    ncx: int
    ncy: int
    ncz: int
    Lx: float = 1.0
    Ly: float = 1.0
    Lz: float = 1.0
    Ra: float = 200.0
    steps: int = 20
    cfl: float = 0.5
    seed: int = 1000
    sigma: float = 0.0          # std of log10 k (0: homogeneous k = 1)
    corr_len: float = 0.1       # correlation length (fraction of Ly)
    modes: int = 1024
    fault: bool = False
    fault_dip_deg: float = 60.0
    fault_mult: float = 100.0
    fault_half_cells: float = 2.5
    noise: float = 0.01
    mu_a: float = 0.0            # viscosity mu = exp(-mu_a T) in the Darcy solve (0: constant)

    @property
    def ncell(self) -> int:
        return self.ncx * self.ncy * self.ncz

    @property
    def h(self) -> tuple[float, float, float]:
        return (self.Lx / self.ncx, self.Ly / self.ncy, self.Lz / self.ncz)

    def scaled(self, ncx: int, ncy: int, ncz: int, **kw) -> "Workload":
        return replace(self, ncx=ncx, ncy=ncy, ncz=ncz, **kw)


# The five benchmark workloads (reference axes).
WORKLOADS = {
    "hrl32": Workload("hrl32: 32x32x32, Ra=200, k=1", 32, 32, 32, Ra=200.0, steps=20, seed=1000),
    "hrl256": Workload("hrl256: 256x128x256 (2x1x2), Ra=500, lognormal k", 256, 128, 256, Lx=2.0, Ly=1.0, Lz=2.0,
                       Ra=500.0, steps=50, seed=1001, sigma=0.5),
    "hrl512_fault": Workload("hrl512_fault: 512x256x512 (2x1x2), Ra=1000, lognormal k + fault", 512, 256, 512,
                             Lx=2.0, Ly=1.0, Lz=2.0, Ra=1000.0, steps=50, seed=1002, sigma=0.5, fault=True),
    "hrl768_fault": Workload("hrl768_fault: 768x384x768 (2x1x2), Ra=2000, mu=exp(-ln10 T), lognormal k + fault", 768,
                             384, 768, Lx=2.0, Ly=1.0, Lz=2.0, Ra=2000.0, steps=10, seed=1004, sigma=0.5, fault=True,
                             mu_a=math.log(10.0)),
}


def rng(seed: int, stream: int = 0) -> np.random.Generator:
    return np.random.Generator(np.random.Philox(key=[seed, stream]))


def initial_temperature(w: Workload) -> np.ndarray:
    """Tc_ex (E array, Fortran order) of the workload."""
    nx, ny, nz = w.ncx, w.ncy, w.ncz
    dy = w.Ly / ny
    yce = (np.arange(ny + 2) - 0.5) * dy          # y of the extended cell centres (ymin = 0)
    T = np.empty((nx + 2, ny + 2, nz + 2), order="F")
    T[...] = (1.0 - yce / w.Ly)[None, :, None]
    u = rng(w.seed, 1).random(nx * ny * nz).reshape((nx, ny, nz), order="F")
    T[1:-1, 1:-1, 1:-1] += w.noise * (2.0 * u - 1.0)
    return T


def lognormal_modes(w: Workload) -> tuple[np.ndarray, np.ndarray]:
    """Wavevectors K (M x 3, C order) and phases phi (M) of the spectral field."""
    g = rng(w.seed, 2)
    ell = w.corr_len * w.Ly
    Z = g.standard_normal((w.modes, 3))
    chi = np.abs(g.standard_normal(w.modes))
    chi = np.maximum(chi, 1e-300)
    K = Z / (ell * chi[:, None])
    phi = g.uniform(0.0, 2.0 * np.pi, w.modes)
    return np.ascontiguousarray(K), phi


def fault_geometry(w: Workload) -> tuple[float, float, np.ndarray, np.ndarray]:
    """(multiplier, half width, normal, point) of the planar fault (multiplier 1: none)."""
    if not w.fault:
        return 1.0, 0.0, np.zeros(3), np.zeros(3)
    dip = math.radians(w.fault_dip_deg)
    n = np.array([math.sin(dip), math.cos(dip), 0.0])
    x0 = np.array([w.Lx / 2, w.Ly / 2, w.Lz / 2])
    h = min(w.h)
    return w.fault_mult, w.fault_half_cells * h, n, x0


def field_args(w: Workload):
    """Arguments of the field generator after (kfv, ncx, ncy, ncz)."""
    K, phi = lognormal_modes(w)
    mult, half, n, x0 = fault_geometry(w)
    dx, dy, dz = w.h
    return (0.0, 0.0, 0.0, dx, dy, dz, K, phi, w.modes, w.sigma, mult, half, n, x0)


def permeability_device(w: Workload, out: DeviceArray) -> DeviceArray:
    """Fill the vertex array kfv on the device (ht3d_lognormal_field)."""
    want = (w.ncx + 1, w.ncy + 1, w.ncz + 1)
    if out.shape != want:
        raise ValueError(f"kfv must have shape {want}")
    if w.sigma == 0.0 and not w.fault:
        out.set(np.ones(want, order="F"))
        return out
    xmin, ymin, zmin, dx, dy, dz, K, phi, M, sigma, mult, half, n, x0 = field_args(w)
    _lib.call("ht3d_lognormal_field", C.c_void_p(out.ptr), w.ncx, w.ncy, w.ncz, C.c_double(xmin), C.c_double(ymin),
              C.c_double(zmin), C.c_double(dx), C.c_double(dy), C.c_double(dz), C.c_void_p(K.ctypes.data),
              C.c_void_p(phi.ctypes.data), int(M), C.c_double(sigma), C.c_double(mult), C.c_double(half),
              C.c_void_p(n.ctypes.data), C.c_void_p(x0.ctypes.data))
    return out


def batch_workloads(n: int = 96, seed: int = 13000, steps: int = 20) -> list[Workload]:
    """Many independent problems of uneven size: edge n from {32,48,64,96,128},
    shape n^3 or n x n/2 x n (vertical halved), Ra log-uniform in [50, 2000],
    sigma(log10 k) uniform in [0, 1]; problem p is seeded with seed + p."""
    g = rng(seed, 0)
    edges = np.array([32, 48, 64, 96, 128])
    out = []
    for p in range(n):
        e = int(g.choice(edges))
        half = bool(g.integers(0, 2))
        ra = float(np.exp(g.uniform(np.log(50.0), np.log(2000.0))))
        sig = float(g.uniform(0.0, 1.0))
        ny = e // 2 if half else e
        out.append(Workload(f"batch{p}", e, ny, e, Lx=1.0, Ly=ny / e, Lz=1.0, Ra=ra, steps=steps, seed=seed + p,
                            sigma=sig))
    return out
