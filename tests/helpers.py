"""Shared helpers of the test suite (staggered shapes, random inputs)."""
import numpy as np


def shapes(nx, ny, nz):
    return {
        "c": (nx, ny, nz), "E": (nx + 2, ny + 2, nz + 2), "X6": (nx + 6, ny + 6, nz + 6),
        "V": (nx + 1, ny + 1, nz + 1), "Fx": (nx + 1, ny, nz), "Fy": (nx, ny + 1, nz), "Fz": (nx, ny, nz + 1),
    }


def rand(shape, rng, lo=-1.0, hi=1.0):
    return np.asfortranarray(rng.uniform(lo, hi, size=shape))


def positive(shape, rng):
    return np.asfortranarray(10.0 ** rng.uniform(-1.0, 1.0, size=shape))


def face_ghost_mask(nx, ny, nz):
    """Interior plus face ghosts of an E array (edges/corners excluded)."""
    m = np.zeros((nx + 2, ny + 2, nz + 2), dtype=bool)
    m[1:-1, 1:-1, 1:-1] = True
    m[[0, -1], 1:-1, 1:-1] = True
    m[1:-1, [0, -1], 1:-1] = True
    m[1:-1, 1:-1, [0, -1]] = True
    return m


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / den)
