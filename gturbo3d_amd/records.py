"""Fortran unformatted sequential records of the restart and snapshot fields.

The reference stores each velocity component as one float32 record of
NX*NY*NZ values, (NX, NY, NZ) column-major, i.e. C order [NZ][NY][NX]
(src/write.f90:12-22, read back by src/read.f90:10-22).  A record is a
4-byte length, the data and the same length again.  A length above
2**31 - 1 does not fit such a marker (a 1024^3 field is 4 GiB), so the
driver splits longer records into subrecords of at most MAXSUB bytes in
gfortran's layout: the leading marker of a subrecord is negative when
another subrecord follows, the trailing marker negative when one precedes.
A record at or under MAXSUB bytes is one plain record, byte for byte the
reference's.
"""
from __future__ import annotations

import numpy as np

MAXSUB = 2147483639  # gfortran's largest subrecord with 4-byte markers


def write_record(path: str, a, maxsub: int = MAXSUB) -> None:
    """One logical record of float32 values (``a`` in C order)."""
    raw = np.ascontiguousarray(a, dtype=np.float32).tobytes()
    n = len(raw)
    with open(path, "wb") as f:
        pos = 0
        while True:
            s = min(maxsub, n - pos)
            lead = -s if pos + s < n else s
            trail = -s if pos > 0 else s
            f.write(np.array([lead], np.int32).tobytes())
            f.write(raw[pos:pos + s])
            f.write(np.array([trail], np.int32).tobytes())
            pos += s
            if pos >= n:
                break


def read_record(path: str, shape) -> np.ndarray:
    """The float32 values of the first logical record of ``path``, reshaped."""
    data = open(path, "rb").read()
    out = bytearray()
    pos = 0
    while True:
        lead = int(np.frombuffer(data[pos:pos + 4], np.int32)[0])
        s = abs(lead)
        out += data[pos + 4:pos + 4 + s]
        trail = int(np.frombuffer(data[pos + 4 + s:pos + 8 + s], np.int32)[0])
        if abs(trail) != s:
            raise ValueError(f"{path}: record markers {lead} / {trail} disagree")
        pos += 8 + s
        if lead >= 0:
            break
    need = int(np.prod(shape)) * 4
    if len(out) != need:
        raise ValueError(f"{path}: record of {len(out)} bytes, expected {need}")
    return np.frombuffer(bytes(out), np.float32).reshape(shape)
