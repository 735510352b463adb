"""On-disk output of the v18 model: VTK rectilinear-grid files and breakpoints.

`write_vtr` writes what WriteVTK's `vtk_grid(name, x, y, z)` + `vtkfile[...] =
...` + `vtk_save` produce for HydroThermal3D_v18.jl:663-687: an XML
RectilinearGrid (`name.vtr`) whose grid points are the cell centres, with one
point-data array per output field (a 3-tuple is a 3-component vector).  The
data are appended raw (little-endian FP64, each block preceded by its byte
count as UInt64), a layout every VTK reader accepts.  `read_vtr` reads such a
file back.

Breakpoints (`Breakpoint%05d.h5`, v18:579-588, restart :384-395) hold Tc_ex,
Pc_ex, time and dt.  `save_breakpoint` writes them as HDF5 in the layout
HDF5.jl gives them (hydrothermal3d_amd.h5io), so the reference's restart code
reads them and a breakpoint of the reference restarts this library's model;
a name ending in `.npz` writes the same four datasets to a NumPy archive
instead (uncompressed, no pickled objects).  `load_breakpoint` reads either.
"""
from __future__ import annotations

import re
import struct
from pathlib import Path

import numpy as np


def _fortran_flat(a: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype="<f8").ravel(order="F"))


def write_vtr(name: str | Path, x, y, z, fields: dict) -> Path:
    """Write `name`.vtr.  x, y, z: grid-point coordinates (1D); fields: name ->
    array of shape (len(x), len(y), len(z)) or a tuple of three such arrays."""
    path = Path(str(name) + ".vtr") if not str(name).endswith(".vtr") else Path(name)
    x, y, z = (np.asarray(c, dtype="<f8").ravel() for c in (x, y, z))
    shape = (len(x), len(y), len(z))
    blocks, head = [], []
    offset = 0

    def add(arr: np.ndarray, name: str, ncomp: int) -> str:
        nonlocal offset
        raw = arr.tobytes()
        blocks.append(struct.pack("<Q", len(raw)) + raw)
        tag = (f'<DataArray type="Float64" Name="{name}" NumberOfComponents="{ncomp}" format="appended" '
               f'offset="{offset}"/>')
        offset += 8 + len(raw)
        return tag

    for fname, val in fields.items():
        if isinstance(val, (tuple, list)):
            comps = [np.asarray(v) for v in val]
            if len(comps) != 3 or any(c.shape != shape for c in comps):
                raise ValueError(f"{fname}: a vector needs three arrays of shape {shape}")
            arr = np.stack([_fortran_flat(c) for c in comps], axis=1).ravel()
            head.append(add(np.ascontiguousarray(arr, dtype="<f8"), fname, 3))
        else:
            if np.shape(val) != shape:
                raise ValueError(f"{fname}: shape {np.shape(val)} != {shape}")
            head.append(add(_fortran_flat(val), fname, 1))
    coords = [add(c, n, 1) for c, n in ((x, "x"), (y, "y"), (z, "z"))]
    ext = f"0 {shape[0] - 1} 0 {shape[1] - 1} 0 {shape[2] - 1}"
    xml = ('<?xml version="1.0"?>\n'
           '<VTKFile type="RectilinearGrid" version="1.0" byte_order="LittleEndian" header_type="UInt64">\n'
           f'  <RectilinearGrid WholeExtent="{ext}">\n'
           f'    <Piece Extent="{ext}">\n'
           '      <PointData>\n' + "".join(f"        {h}\n" for h in head) + '      </PointData>\n'
           '      <CellData>\n      </CellData>\n'
           '      <Coordinates>\n' + "".join(f"        {c}\n" for c in coords) + '      </Coordinates>\n'
           '    </Piece>\n  </RectilinearGrid>\n  <AppendedData encoding="raw">\n   _')
    with open(path, "wb") as fh:
        fh.write(xml.encode())
        for b in blocks:
            fh.write(b)
        fh.write(b"\n  </AppendedData>\n</VTKFile>\n")
    return path


def read_vtr(path: str | Path) -> tuple[tuple[np.ndarray, np.ndarray, np.ndarray], dict]:
    """Read a file written by write_vtr: ((x, y, z), {name: array or 3-tuple})."""
    data = Path(path).read_bytes()
    cut = data.index(b'<AppendedData encoding="raw">')
    start = data.index(b"_", cut) + 1
    head = data[:cut].decode()
    ext = [int(v) for v in re.search(r'WholeExtent="([^"]+)"', head).group(1).split()]
    shape = (ext[1] + 1, ext[3] + 1, ext[5] + 1)
    out, coords = {}, {}
    for m in re.finditer(r'<DataArray type="Float64" Name="([^"]+)" NumberOfComponents="(\d)" '
                         r'format="appended" offset="(\d+)"/>', head):
        name, ncomp, off = m.group(1), int(m.group(2)), int(m.group(3))
        (nbytes,) = struct.unpack("<Q", data[start + off:start + off + 8])
        arr = np.frombuffer(data[start + off + 8:start + off + 8 + nbytes], dtype="<f8")
        if name in ("x", "y", "z"):
            coords[name] = arr.copy()
        elif ncomp == 3:
            comp = arr.reshape(-1, 3)
            out[name] = tuple(np.asfortranarray(comp[:, q].reshape(shape, order="F")) for q in range(3))
        else:
            out[name] = np.asfortranarray(arr.reshape(shape, order="F"))
    return (coords["x"], coords["y"], coords["z"]), out


def save_breakpoint(name: str | Path, Tc_ex: np.ndarray, Pc_ex: np.ndarray, time: float, dt: float) -> Path:
    """Breakpoint%05d (v18:579-588): Tc_ex, Pc_ex, time, dt, as `name`.h5
    (HDF5, the reference's format) or, for a name ending in .npz, a NumPy
    archive."""
    if str(name).endswith(".npz"):
        path = Path(name)
        np.savez(path, Tc_ex=np.asfortranarray(Tc_ex), Pc_ex=np.asfortranarray(Pc_ex), time=np.float64(time),
                 dt=np.float64(dt))
        return path
    from .h5io import write_datasets
    path = Path(name) if str(name).endswith(".h5") else Path(str(name) + ".h5")
    return write_datasets(path, {"Tc_ex": Tc_ex, "Pc_ex": Pc_ex, "time": float(time), "dt": float(dt)})


def load_breakpoint(path: str | Path) -> dict:
    """The datasets of a breakpoint (restart, v18:384-395): an HDF5 file as the
    reference writes it, or a NumPy archive (.npz)."""
    if str(path).endswith(".npz"):
        with np.load(path, allow_pickle=False) as f:
            return {"Tc_ex": np.asfortranarray(f["Tc_ex"]), "Pc_ex": np.asfortranarray(f["Pc_ex"]),
                    "time": float(f["time"]), "dt": float(f["dt"])}
    from .h5io import read_datasets
    return read_datasets(path, ("Tc_ex", "Pc_ex", "time", "dt"))
