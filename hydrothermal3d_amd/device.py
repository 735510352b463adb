"""Device arrays for the HIP path (the `Data.Array` of the reference drivers).

A DeviceArray is an FP64 (or FP32) column-major buffer in HBM owned through
ht3d_alloc/ht3d_free.  Shapes follow Julia's (n1, n2, n3) with n1 fastest, so
`DeviceArray.from_host(A).get()` round-trips a Fortran-ordered NumPy array bit
for bit (HydroThermal3D_v7.jl:153-157 does the same with Data.Array / Array).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


class DeviceArray:
    __slots__ = ("ptr", "shape", "dtype", "nbytes", "_owner", "__weakref__")

    def __init__(self, shape, dtype=np.float64, *, ptr: int | None = None, owner=None):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self._owner = owner
        if ptr is None:
            out = C.c_void_p()
            _lib.call("ht3d_alloc", C.c_size_t(self.nbytes), C.byref(out))
            self.ptr = int(out.value or 0)
        else:
            self.ptr = int(ptr)

    # --------------------------------------------------------------- factories
    @classmethod
    def zeros(cls, shape, dtype=np.float64) -> "DeviceArray":
        a = cls(shape, dtype)
        a.fill_zero()
        return a

    @classmethod
    def ones(cls, shape) -> "DeviceArray":
        return cls.from_host(np.ones(shape, order="F"))

    @classmethod
    def from_host(cls, host: np.ndarray, dtype=None) -> "DeviceArray":
        host = np.asarray(host, dtype=dtype or host.dtype)
        a = cls(host.shape, host.dtype)
        a.set(host)
        return a

    # --------------------------------------------------------------- transfers
    def set(self, host: np.ndarray) -> "DeviceArray":
        host = np.asfortranarray(np.asarray(host, dtype=self.dtype))
        if host.shape != self.shape:
            raise ValueError(f"shape mismatch: device {self.shape}, host {host.shape}")
        _lib.call("ht3d_h2d", C.c_void_p(self.ptr), C.c_void_p(host.ctypes.data), C.c_size_t(self.nbytes))
        return self

    def get(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype, order="F")
        _lib.call("ht3d_d2h", C.c_void_p(out.ctypes.data), C.c_void_p(self.ptr), C.c_size_t(self.nbytes))
        return out

    def fill_zero(self) -> "DeviceArray":
        _lib.call("ht3d_memset", C.c_void_p(self.ptr), 0, C.c_size_t(self.nbytes))
        return self

    def copy_from(self, other: "DeviceArray") -> "DeviceArray":
        if other.shape != self.shape or other.dtype != self.dtype:
            raise ValueError("copy_from: shape/dtype mismatch")
        _lib.call("ht3d_d2d", C.c_void_p(self.ptr), C.c_void_p(other.ptr), C.c_size_t(self.nbytes))
        return self

    # --------------------------------------------------------------- misc
    @property
    def size(self) -> int:
        return int(np.prod(self.shape))

    def __repr__(self) -> str:
        return f"DeviceArray(shape={self.shape}, dtype={self.dtype.name}, ptr=0x{self.ptr:x})"

    def free(self) -> None:
        if self.ptr and self._owner is None and _lib._lib is not None:
            try:
                _lib.lib().ht3d_free(C.c_void_p(self.ptr))
            except Exception:
                pass
        self.ptr = 0

    def __del__(self):
        self.free()


def reduce(a: DeviceArray, op: int) -> float:
    """maximum_g / minimum_g / sum_g over a device array (tools/Macros.jl:20-25)."""
    out = C.c_double()
    _lib.call("ht3d_reduce", C.byref(out), C.c_void_p(a.ptr), C.c_size_t(a.size), op)
    return out.value


def maximum_g(a: DeviceArray) -> float:
    return reduce(a, _lib.MAX)


def minimum_g(a: DeviceArray) -> float:
    return reduce(a, _lib.MIN)


def sum_g(a: DeviceArray) -> float:
    return reduce(a, _lib.SUM)


def mean_abs_g(a: DeviceArray) -> float:
    """mean_g(abs.(A)) (HydroThermal3D_v7.jl:180) without the broadcast temporary."""
    return reduce(a, _lib.SUM_ABS) / a.size


def max_abs_g(a: DeviceArray) -> float:
    """maximum_g(abs.(A)) (HydroThermal3D_v7.jl:220)."""
    return reduce(a, _lib.MAX_ABS)
