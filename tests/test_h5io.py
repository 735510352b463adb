"""HDF5 files in the reference's layout (hydrothermal3d_amd.h5io), CPU only.

Breakpoints (Dimensional/HydroThermal3D_v18.jl:579-588, read back at :384-395)
and Nu(t) files (Dimensionless/HydroThermal3D_Alban_v6.jl:500-505) as Julia's
HDF5.jl writes them: a column-major (n1, n2, n3) array is a dataset of
dataspace (n3, n2, n1) over the same bytes, a scalar a scalar dataspace, all
H5T_IEEE_F64LE.  The files are checked with the HDF5 project's own h5dump, not
only read back with this library's reader.
"""
import re
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from hydrothermal3d_amd import h5io
from hydrothermal3d_amd.vtkio import load_breakpoint, save_breakpoint


def _have_hdf5() -> bool:
    try:
        h5io.lib()
        return True
    except RuntimeError:
        return False


pytestmark = pytest.mark.skipif(not _have_hdf5(), reason="no HDF5 C library in this image")


def _h5dump() -> str | None:
    for cand in (shutil.which("h5dump"), "/opt/conda/bin/h5dump"):
        if cand and Path(cand).exists():
            return cand
    return None


def _dump_values(path, name):
    """Values of dataset `name` in the order h5dump prints them (C order of
    the dataspace), at full precision."""
    out = subprocess.run([_h5dump(), "-d", name, "-w", "0", "-m", "%.17g", str(path)], capture_output=True, text=True,
                         check=True).stdout
    body = out[out.index("DATA {") + len("DATA {"):]
    body = re.sub(r"\(\d+(,\d+)*\):", " ", body)
    return np.array([float(t) for t in re.findall(r"-?\d[\d.eE+-]*|-?inf|nan", body.replace(",", " "))])


def test_breakpoint_hdf5_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    T = np.asfortranarray(rng.standard_normal((7, 5, 4)))
    P = np.asfortranarray(rng.standard_normal((7, 5, 4)))
    p = save_breakpoint(tmp_path / "Breakpoint00012", T, P, 2.5e-3, 1.0 / 3.0)
    assert p.name == "Breakpoint00012.h5"
    b = load_breakpoint(p)
    assert np.array_equal(b["Tc_ex"], T) and np.array_equal(b["Pc_ex"], P)
    assert b["Tc_ex"].flags.f_contiguous and b["Tc_ex"].shape == (7, 5, 4)
    assert b["time"] == 2.5e-3 and b["dt"] == 1.0 / 3.0


@pytest.mark.skipif(_h5dump() is None, reason="h5dump not installed")
def test_breakpoint_layout_as_hdf5jl_writes_it(tmp_path):
    rng = np.random.default_rng(4)
    T = np.asfortranarray(rng.standard_normal((6, 5, 4)))
    P = np.asfortranarray(rng.standard_normal((6, 5, 4)))
    p = save_breakpoint(tmp_path / "Breakpoint00001", T, P, 1.25, 0.5)
    head = subprocess.run([_h5dump(), "-H", str(p)], capture_output=True, text=True, check=True).stdout
    for name in ("Tc_ex", "Pc_ex"):
        block = head[head.index(f'DATASET "{name}"'):]
        assert "H5T_IEEE_F64LE" in block[:200]
        assert "SIMPLE { ( 4, 5, 6 ) / ( 4, 5, 6 ) }" in block[:200], block[:200]   # extents reversed
    for name in ("time", "dt"):
        block = head[head.index(f'DATASET "{name}"'):]
        assert "H5T_IEEE_F64LE" in block[:200] and "SCALAR" in block[:200]
    # the bytes are column-major: h5dump's C-order walk over (4, 5, 6) is T's Fortran order
    assert np.array_equal(_dump_values(p, "/Tc_ex"), T.ravel(order="F"))
    assert np.array_equal(_dump_values(p, "/Pc_ex"), P.ravel(order="F"))
    assert _dump_values(p, "/time").tolist() == [1.25] and _dump_values(p, "/dt").tolist() == [0.5]


@pytest.mark.skipif(_h5dump() is None, reason="h5dump not installed")
def test_nu_history_file(tmp_path):
    t = np.linspace(0.0, 1e-3, 9)
    nu = 1.0 + np.sqrt(t)
    p = h5io.write_nu_history(tmp_path / "HT3DOutput00009.h5", t, nu)
    head = subprocess.run([_h5dump(), "-H", str(p)], capture_output=True, text=True, check=True).stdout
    assert head.count("SIMPLE { ( 9 ) / ( 9 ) }") == 2
    assert np.array_equal(_dump_values(p, "/time"), t) and np.array_equal(_dump_values(p, "/Nu"), nu)
    r = h5io.read_datasets(p, ("time", "Nu"))
    assert np.array_equal(r["time"], t) and np.array_equal(r["Nu"], nu)


def test_npz_breakpoints_still_read(tmp_path):
    rng = np.random.default_rng(5)
    T = np.asfortranarray(rng.standard_normal((4, 4, 3)))
    p = save_breakpoint(tmp_path / "Breakpoint00002.npz", T, -T, 1.0, 0.25)
    b = load_breakpoint(p)
    assert np.array_equal(b["Tc_ex"], T) and np.array_equal(b["Pc_ex"], -T) and b["dt"] == 0.25


def test_missing_dataset_raises(tmp_path):
    p = h5io.write_datasets(tmp_path / "x.h5", {"a": np.arange(3.0)})
    with pytest.raises(OSError):
        h5io.read_datasets(p, ("b",))
