"""Reader for Fortran sequential unformatted files (4-byte record markers),
the record format that gfortran/flang/amdflang write by default."""
import struct

import numpy as np


def read_records(path):
    recs = []
    with open(path, "rb") as f:
        data = f.read()
    pos = 0
    while pos < len(data):
        (n,) = struct.unpack_from("<i", data, pos)
        pos += 4
        recs.append(data[pos:pos + n])
        pos += n
        (n2,) = struct.unpack_from("<i", data, pos)
        if n2 != n:
            raise ValueError(f"record marker mismatch at byte {pos}: {n} vs {n2}")
        pos += 4
    return recs


def ints(rec):
    return np.frombuffer(rec, dtype="<i4")


def doubles(rec):
    return np.frombuffer(rec, dtype="<f8")
