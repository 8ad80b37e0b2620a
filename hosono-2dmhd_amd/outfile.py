"""Reader of the Fortran driver's output files (fortran/hmhd_out.f90).

Format: magic b'HMHDOUT1', int32 record count, then per record a 16-byte
blank-padded name, int32 type (1 = int32, 2 = float64), int32 rank, int32
dims[3] and the data in Fortran order, all little-endian.  Two-dimensional
fields dims (nx, ny) come back as numpy arrays of shape (ny, nx) (element
[j, i] is field(i+1, j+1)); `cons` (8, nx, ny) comes back as (ny, nx, 8).
"""
from __future__ import annotations

import pathlib

import numpy as np

MAGIC = b"HMHDOUT1"


def read_output(path) -> dict:
    raw = pathlib.Path(path).read_bytes()
    if raw[:8] != MAGIC:
        raise ValueError(f"{path}: not an hmhd output file")
    n = int(np.frombuffer(raw, "<i4", 1, 8)[0])
    off = 12
    out = {}
    for _ in range(n):
        name = raw[off:off + 16].decode().strip()
        typ, rank, d1, d2, d3 = (int(x) for x in np.frombuffer(raw, "<i4", 5, off + 16))
        off += 36
        dt = np.dtype("<i4") if typ == 1 else np.dtype("<f8")
        count = d1 * d2 * d3
        a = np.frombuffer(raw, dt, count, off).copy()
        off += count * dt.itemsize
        if rank == 1 and count == 1:
            out[name] = a[0].item()
        elif rank == 2:
            out[name] = a.reshape(d2, d1)
        else:
            out[name] = a.reshape(d3, d2, d1)
    return out
