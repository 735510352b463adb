"""VTK rectilinear-grid output and breakpoints (hydrothermal3d_amd.vtkio), CPU only."""
import numpy as np

from hydrothermal3d_amd.vtkio import load_breakpoint, read_vtr, save_breakpoint, write_vtr


def test_vtr_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    shape = (5, 4, 3)
    x, y, z = np.linspace(0, 1, 5), np.linspace(-2, 0, 4), np.array([0.1, 0.4, 0.9])
    T = np.asfortranarray(rng.standard_normal(shape))
    T[1, 2, 0] = np.nan
    V = tuple(np.asfortranarray(rng.standard_normal(shape)) for _ in range(3))
    path = write_vtr(tmp_path / "HT3DOutput00010", x, y, z, {"Temperature": T, "Velocity": V})
    assert path.name == "HT3DOutput00010.vtr"
    head = path.read_bytes()[:600].decode(errors="replace")
    assert 'type="RectilinearGrid"' in head and 'WholeExtent="0 4 0 3 0 2"' in head
    assert 'Name="Velocity" NumberOfComponents="3"' in head
    (xr, yr, zr), f = read_vtr(path)
    assert np.array_equal(xr, x) and np.array_equal(yr, y) and np.array_equal(zr, z)
    assert np.array_equal(f["Temperature"], T, equal_nan=True)
    for a, b in zip(f["Velocity"], V):
        assert np.array_equal(a, b)
    # the vector is stored interleaved, point by point, x fastest
    raw = path.read_bytes()
    start = raw.index(b"_", raw.index(b"<AppendedData")) + 1
    n = int(np.prod(shape))
    off = 8 + 8 * n + 8
    first = np.frombuffer(raw[start + off:start + off + 24], dtype="<f8")
    assert np.array_equal(first, [V[0][0, 0, 0], V[1][0, 0, 0], V[2][0, 0, 0]])


def test_breakpoint_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    T = np.asfortranarray(rng.standard_normal((6, 5, 4)))
    P = np.asfortranarray(rng.standard_normal((6, 5, 4)))
    p = save_breakpoint(tmp_path / "Breakpoint00003", T, P, 1.25, 0.5)
    b = load_breakpoint(p)
    assert np.array_equal(b["Tc_ex"], T) and np.array_equal(b["Pc_ex"], P)
    assert b["time"] == 1.25 and b["dt"] == 0.5 and b["Tc_ex"].flags.f_contiguous
