"""Matrix::D3D / D2D .dat files (Glioma/Matrix.h:26-34, 593-665).

Layout: int32 header {1234, dim, nx, ny[, nz]}, int32 type id (0 double,
1 float, 2 int), then the values with x fastest.  3D arrays are returned and
taken as numpy arrays of shape (nz, ny, nx); 2D as (ny, nx), so ROI.dat's
N x 3 point list reads as shape (3, N).
"""
from __future__ import annotations

import numpy as np

_TYPES = {0: np.float64, 1: np.float32, 2: np.int32}


def read_dat(path: str) -> np.ndarray:
    with open(path, "rb") as f:
        head = np.frombuffer(f.read(8), np.int32)
        if head.size != 2 or head[0] != 1234 or head[1] not in (2, 3):
            raise ValueError("%s: not a Matrix .dat file" % path)
        dim = int(head[1])
        dims = np.frombuffer(f.read(4 * dim), np.int32)
        tid = int(np.frombuffer(f.read(4), np.int32)[0])
        if tid not in _TYPES:
            raise ValueError("%s: unknown type id %d" % (path, tid))
        n = int(np.prod(dims.astype(np.int64)))
        data = np.frombuffer(f.read(n * np.dtype(_TYPES[tid]).itemsize), _TYPES[tid])
        if data.size != n:
            raise ValueError("%s: truncated" % path)
    return data.astype(np.float64).reshape(tuple(int(d) for d in dims[::-1]))


def write_dat(path: str, a: np.ndarray) -> None:
    """Write a 2D (ny, nx) or 3D (nz, ny, nx) array as doubles."""
    a = np.ascontiguousarray(a, np.float64)
    if a.ndim not in (2, 3):
        raise ValueError("write_dat: 2D or 3D arrays only")
    head = np.array([1234, a.ndim] + list(a.shape[::-1]) + [0], np.int32)
    with open(path, "wb") as f:
        f.write(head.tobytes())
        f.write(a.tobytes())
