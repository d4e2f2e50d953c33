"""The reference's file formats (pqdts.f90:1843-1876, 2146-2160, 2183-2197;
pqdts.py:40-58, 94-107).

Sequential unformatted Fortran files with 4-byte record markers:
  [int64 M] [int64 N] [int64 D] [int64 localrows] [int32 rank] [float64 X(localrows, N), column-major]
"""
from __future__ import annotations

import struct

import numpy as np


def _records(buf: bytes):
    pos = 0
    while pos < len(buf):
        (n,) = struct.unpack_from("<i", buf, pos)
        pos += 4
        yield buf[pos:pos + n]
        pos += n
        (n2,) = struct.unpack_from("<i", buf, pos)
        if n2 != n:
            raise ValueError("corrupt Fortran record markers")
        pos += 4


def read_state(path: str):
    """Returns (M, N, D, localrows, rank, X) with X as an (M, N) array."""
    with open(path, "rb") as f:
        recs = list(_records(f.read()))
    if len(recs) != 6:
        raise ValueError(f"{path}: expected 6 records, found {len(recs)}")
    ints = []
    for r in recs[:5]:
        ints.append(int(np.frombuffer(r, dtype="<i8" if len(r) == 8 else "<i4")[0]))
    M, N, D, lr, rank = ints
    X = np.frombuffer(recs[5], dtype="<f8").reshape((lr, N), order="F")
    return M, N, D, lr, rank, np.array(X)


def write_state(path: str, M: int, N: int, D: int, X: np.ndarray, rank: int = 0, rank_bytes: int = 4):
    """Writes a state file; rank_bytes=4 as the solver writes, 8 as pqdts.py writes."""
    X = np.asarray(X, dtype="<f8")
    lr = X.shape[0]
    out = bytearray()

    def rec(b: bytes):
        out.extend(struct.pack("<i", len(b)))
        out.extend(b)
        out.extend(struct.pack("<i", len(b)))

    for v in (M, N, D, lr):
        rec(struct.pack("<q", v))
    rec(struct.pack("<i" if rank_bytes == 4 else "<q", rank))
    rec(np.asfortranarray(X).tobytes(order="F"))
    with open(path, "wb") as f:
        f.write(bytes(out))
