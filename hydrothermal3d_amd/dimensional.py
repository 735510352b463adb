"""Characteristic scales of the dimensional (v18) model.

`Scaling` and `scale_me` mirror the `scaling` struct and `scale_me!` of
Dimensional/kernels_HT3D_dimensional.jl:3-36: four primary scales (T, V, L, eta)
set by the driver, and the derived ones (time, stress, mass, density, heat
capacity, thermal conductivity, permeability ...) that the property kernels
take as arguments.  `v18_scaling()` returns the values of
Dimensional/HydroThermal3D_v18.jl:137-142.
"""
from __future__ import annotations

from dataclasses import dataclass


@dataclass
class Scaling:
    T: float = float("nan")
    V: float = float("nan")
    L: float = float("nan")
    eta: float = float("nan")
    t: float = float("nan")
    a: float = float("nan")
    eps: float = float("nan")
    sigma: float = float("nan")
    m: float = float("nan")
    rho: float = float("nan")
    F: float = float("nan")
    J: float = float("nan")
    W: float = float("nan")
    C: float = float("nan")
    kt: float = float("nan")
    kf: float = float("nan")


def scale_me(s: Scaling) -> Scaling:
    """Derived scales, in the order and with the formulas of scale_me!
    (x^2 = x*x and x^3 = x*x*x, as Julia evaluates literal powers)."""
    s.t = s.L / s.V
    s.a = s.V / s.t
    s.eps = 1.0 / s.t
    s.sigma = s.eta / s.t
    s.m = s.sigma * s.L * (s.t * s.t)
    s.rho = s.m / (s.L * s.L * s.L)
    s.F = s.m * s.L / (s.t * s.t)
    s.J = s.m * (s.L * s.L) / (s.t * s.t)
    s.W = s.J / s.t
    s.C = s.J / s.m / s.T
    s.kt = s.W / s.L / s.T
    s.kf = s.L * s.L
    return s


def v18_scaling() -> Scaling:
    return scale_me(Scaling(T=500.0, V=1e-9, L=700.0, eta=1e10))
