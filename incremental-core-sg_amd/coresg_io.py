"""The `.coresg` on-disk format, read and written byte-compatibly.

Layout (reference src/io.cpp:195-261): magic "CORESG01"; u32 k_max, dim,
stride, metric, missing-edge mode; u64 n; u64 float count then the padded rows;
four u64-counted vectors ids (u32), dist (f32), flags (u8), count (u32); a
u64-counted vector of 12-byte bridge records {u32 u, u32 v, f32 w}; u64 repair
edge count. Little-endian, no padding between fields.
"""
from __future__ import annotations

import dataclasses
import struct

import numpy as np

MAGIC = b"CORESG01"


@dataclasses.dataclass
class CoreSGFile:
    k_max: int
    dim: int
    stride: int
    metric: int
    mode: int
    points: np.ndarray   # (n, stride) float32, zero padded
    ids: np.ndarray      # (n*k,) uint32
    dist: np.ndarray     # (n*k,) float32
    flags: np.ndarray    # (n*k,) uint8
    count: np.ndarray    # (n,) uint32
    bridges: np.ndarray  # (m, 3) uint32 records; column 2 is the float32 bits of w
    repair_count: int

    @property
    def n(self) -> int:
        return int(self.count.size)

    def bridge_uvw(self):
        b = self.bridges.reshape(-1, 3)
        return b[:, 0].copy(), b[:, 1].copy(), b[:, 2].copy().view(np.float32)


def _vec(buf: bytes, pos: int, dtype):
    (count,) = struct.unpack_from("<Q", buf, pos)
    pos += 8
    dt = np.dtype(dtype)
    arr = np.frombuffer(buf, dt, count, pos).copy()
    return arr, pos + count * dt.itemsize


def loads(buf: bytes) -> CoreSGFile:
    if buf[:8] != MAGIC:
        raise ValueError("not a coresg file")
    k_max, dim, stride, metric, mode = struct.unpack_from("<5I", buf, 8)
    n, floats = struct.unpack_from("<QQ", buf, 28)
    pos = 44
    if floats != n * stride:
        raise ValueError("inconsistent coresg header")
    pts = np.frombuffer(buf, np.float32, floats, pos).reshape(n, stride).copy()
    pos += floats * 4
    ids, pos = _vec(buf, pos, np.uint32)
    dist, pos = _vec(buf, pos, np.float32)
    flags, pos = _vec(buf, pos, np.uint8)
    count, pos = _vec(buf, pos, np.uint32)
    (nb,) = struct.unpack_from("<Q", buf, pos)
    pos += 8
    bridges = np.frombuffer(buf, np.uint32, nb * 3, pos).reshape(nb, 3).copy()
    pos += nb * 12
    (repair,) = struct.unpack_from("<Q", buf, pos)
    pos += 8
    if pos != len(buf):
        raise ValueError("trailing bytes in coresg file")
    return CoreSGFile(k_max, dim, stride, metric, mode, pts, ids, dist, flags, count, bridges,
                      int(repair))


def load(path: str) -> CoreSGFile:
    with open(path, "rb") as f:
        return loads(f.read())


def dumps(f: CoreSGFile) -> bytes:
    parts = [MAGIC, struct.pack("<5I", f.k_max, f.dim, f.stride, f.metric, f.mode),
             struct.pack("<QQ", f.points.shape[0], f.points.size),
             np.ascontiguousarray(f.points, np.float32).tobytes()]
    for arr, dt in ((f.ids, np.uint32), (f.dist, np.float32), (f.flags, np.uint8),
                    (f.count, np.uint32)):
        a = np.ascontiguousarray(arr, dt)
        parts += [struct.pack("<Q", a.size), a.tobytes()]
    b = np.ascontiguousarray(f.bridges, np.uint32).reshape(-1, 3)
    parts += [struct.pack("<Q", b.shape[0]), b.tobytes(), struct.pack("<Q", f.repair_count)]
    return b"".join(parts)


def save(path: str, f: CoreSGFile) -> None:
    with open(path, "wb") as fh:
        fh.write(dumps(f))


def bridges_from_uvw(u, v, w) -> np.ndarray:
    rec = np.empty((len(u), 3), np.uint32)
    rec[:, 0] = u
    rec[:, 1] = v
    rec[:, 2] = np.ascontiguousarray(w, np.float32).view(np.uint32)
    return rec
