"""HDF5 files as the reference's drivers write them, through the HDF5 C library.

The reference writes, with Julia's HDF5.jl:
* breakpoints `Breakpoint%05d.h5` with datasets Tc_ex, Pc_ex (3-D), time and
  dt (scalars) (Dimensional/HydroThermal3D_v18.jl:579-588), read back on
  restart (:384-395);
* Nu(t) files `HT3DOutput%05d.h5` with datasets time and Nu (1-D)
  (Dimensionless/HydroThermal3D_Alban_v6.jl:500-505).

HDF5.jl stores a Julia (column-major) array of size (n1, n2, n3) as a dataset
whose dataspace lists the extents reversed, (n3, n2, n1), over the same bytes;
a Julia scalar becomes a scalar dataspace.  Both are FP64, which on disk is
H5T_IEEE_F64LE.  `write_datasets` and `read_datasets` follow these rules, so
a driver that swaps to this library reads and writes the same files.

The HDF5 shared library (1.10 or later) is found through HT3D_HDF5_LIB, then
the loader's search path, then /opt/conda/lib; without it these calls raise.
"""
from __future__ import annotations

import ctypes as C
import ctypes.util
import os
from pathlib import Path

import numpy as np

hid_t = C.c_int64
herr_t = C.c_int
hsize_t = C.c_ulonglong

H5F_ACC_RDONLY = 0x0000
H5F_ACC_TRUNC = 0x0002
H5P_DEFAULT = 0
H5S_ALL = 0
H5S_SCALAR = 0

_L = None


def _candidates():
    env = os.environ.get("HT3D_HDF5_LIB")
    if env:
        yield env
    found = ctypes.util.find_library("hdf5")
    if found:
        yield found
    for d in ("/opt/conda/lib", "/usr/lib/x86_64-linux-gnu/hdf5/serial", "/usr/lib/x86_64-linux-gnu", "/usr/lib64"):
        for name in ("libhdf5.so", "libhdf5.so.103", "libhdf5_serial.so"):
            p = Path(d) / name
            if p.exists():
                yield str(p)


def lib() -> C.CDLL:
    """The HDF5 C library, initialised (H5open) and with its error stack
    printing switched off (failures come back as negative return values)."""
    global _L
    if _L is not None:
        return _L
    errors = []
    for cand in _candidates():
        try:
            L = C.CDLL(cand)
            break
        except OSError as e:
            errors.append(f"{cand}: {e}")
    else:
        raise RuntimeError("the HDF5 C library (libhdf5) was not found; set HT3D_HDF5_LIB to its path"
                           + (f" ({'; '.join(errors)})" if errors else ""))
    for name, res, args in (
            ("H5open", herr_t, []),
            ("H5Eset_auto2", herr_t, [hid_t, C.c_void_p, C.c_void_p]),
            ("H5Fcreate", hid_t, [C.c_char_p, C.c_uint, hid_t, hid_t]),
            ("H5Fopen", hid_t, [C.c_char_p, C.c_uint, hid_t]),
            ("H5Fclose", herr_t, [hid_t]),
            ("H5Screate", hid_t, [C.c_int]),
            ("H5Screate_simple", hid_t, [C.c_int, C.POINTER(hsize_t), C.POINTER(hsize_t)]),
            ("H5Sget_simple_extent_ndims", C.c_int, [hid_t]),
            ("H5Sget_simple_extent_dims", C.c_int, [hid_t, C.POINTER(hsize_t), C.POINTER(hsize_t)]),
            ("H5Sclose", herr_t, [hid_t]),
            ("H5Dcreate2", hid_t, [hid_t, C.c_char_p, hid_t, hid_t, hid_t, hid_t, hid_t]),
            ("H5Dopen2", hid_t, [hid_t, C.c_char_p, hid_t]),
            ("H5Dget_space", hid_t, [hid_t]),
            ("H5Dwrite", herr_t, [hid_t, hid_t, hid_t, hid_t, hid_t, C.c_void_p]),
            ("H5Dread", herr_t, [hid_t, hid_t, hid_t, hid_t, hid_t, C.c_void_p]),
            ("H5Dclose", herr_t, [hid_t])):
        f = getattr(L, name)
        f.restype, f.argtypes = res, args
    if L.H5open() < 0:
        raise RuntimeError("H5open failed")
    L.H5Eset_auto2(0, None, None)   # H5E_DEFAULT: no automatic error printing
    _L = L
    return L


def _f64(L) -> int:
    """H5T_NATIVE_DOUBLE (a library global, valid after H5open)."""
    return hid_t.in_dll(L, "H5T_NATIVE_DOUBLE_g").value


def _ok(rc: int, what: str) -> int:
    if rc < 0:
        raise OSError(f"HDF5: {what} failed")
    return rc


def write_datasets(path: str | Path, data: dict) -> Path:
    """Create (or truncate) the HDF5 file `path` with one FP64 dataset per
    entry: arrays in HDF5.jl's layout (column-major bytes, extents reversed),
    Python/NumPy scalars as scalar datasets."""
    L = lib()
    path = Path(path)
    f = _ok(L.H5Fcreate(str(path).encode(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), f"creating {path}")
    try:
        t = _f64(L)
        for name, val in data.items():
            a = np.asarray(val, dtype=np.float64)
            if a.ndim == 0:
                s = _ok(L.H5Screate(H5S_SCALAR), "H5Screate")
                buf = np.array([float(a)], dtype=np.float64)
            else:
                dims = (hsize_t * a.ndim)(*reversed(a.shape))
                s = _ok(L.H5Screate_simple(a.ndim, dims, None), "H5Screate_simple")
                buf = np.ascontiguousarray(a.ravel(order="F"))
            try:
                d = _ok(L.H5Dcreate2(f, name.encode(), t, s, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                        f"creating dataset {name}")
                try:
                    _ok(L.H5Dwrite(d, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.ctypes.data), f"writing {name}")
                finally:
                    L.H5Dclose(d)
            finally:
                L.H5Sclose(s)
    finally:
        _ok(L.H5Fclose(f), f"closing {path}")
    return path


def read_datasets(path: str | Path, names) -> dict:
    """Read FP64 datasets of an HDF5 file: arrays come back Fortran-ordered
    with the extents HDF5.jl gives them (the dataspace's, reversed); scalar
    datasets come back as float."""
    L = lib()
    path = Path(path)
    f = _ok(L.H5Fopen(str(path).encode(), H5F_ACC_RDONLY, H5P_DEFAULT), f"opening {path}")
    out = {}
    try:
        t = _f64(L)
        for name in names:
            d = _ok(L.H5Dopen2(f, name.encode(), H5P_DEFAULT), f"opening dataset {name} of {path}")
            try:
                s = _ok(L.H5Dget_space(d), "H5Dget_space")
                try:
                    nd = _ok(L.H5Sget_simple_extent_ndims(s), "H5Sget_simple_extent_ndims")
                    dims = (hsize_t * max(nd, 1))()
                    if nd > 0:
                        _ok(L.H5Sget_simple_extent_dims(s, dims, None), "H5Sget_simple_extent_dims")
                    shape = tuple(int(dims[q]) for q in reversed(range(nd)))
                finally:
                    L.H5Sclose(s)
                buf = np.empty(int(np.prod(shape)) if nd else 1, dtype=np.float64)
                _ok(L.H5Dread(d, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.ctypes.data), f"reading {name}")
            finally:
                L.H5Dclose(d)
            out[name] = float(buf[0]) if nd == 0 else np.asfortranarray(buf.reshape(shape, order="F"))
    finally:
        _ok(L.H5Fclose(f), f"closing {path}")
    return out


def write_nu_history(path: str | Path, time, nu) -> Path:
    """HT3DOutput%05d.h5 of Dimensionless/HydroThermal3D_Alban_v6.jl:500-505:
    the time and Nu(t) vectors of a run so far."""
    return write_datasets(path, {"time": np.asarray(time, dtype=np.float64).ravel(),
                                 "Nu": np.asarray(nu, dtype=np.float64).ravel()})
