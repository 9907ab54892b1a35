"""Reader for the tagged-array files written by oracle/ref_dump.cpp."""
from __future__ import annotations

import numpy as np

_DTYPES = {1: np.uint8, 4: np.uint32, 5: np.int32, 8: np.uint64, 9: np.float32, 10: np.float64}


def read_tagged(path: str) -> dict:
    out = {}
    with open(path, "rb") as f:
        buf = f.read()
    pos = 0
    while pos < len(buf):
        name = buf[pos:pos + 32].split(b"\0", 1)[0].decode()
        dtype = int(np.frombuffer(buf, np.uint32, 1, pos + 32)[0])
        count = int(np.frombuffer(buf, np.uint64, 1, pos + 36)[0])
        pos += 44
        dt = np.dtype(_DTYPES[dtype])
        out[name] = np.frombuffer(buf, dt, count, pos).copy()
        pos += count * dt.itemsize
    return out
