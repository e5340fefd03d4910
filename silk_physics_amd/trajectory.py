"""Reader of the trajectory files silk_headless writes (.silktraj).

Layout (little endian):
  magic        8 bytes  "SILKTRJ1"
  objects      u32
  per object   kind u32 (0 cloth, 1 obstacle), moving u32 (1 if its positions
               are stored per frame: cloths and rotating obstacles), name
               length u32 and UTF-8 bytes, nv u32, nf u32, positions at load
               (after the scene transform) f32[3 nv], faces i32[3 nf]
  frames       u32
  per frame    time f32, then f32[3 nv] for every moving object in order
Frame 0 is the state before the first step; frame k the state after step k.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field

import numpy as np


@dataclass
class TrajectoryObject:
    name: str
    cloth: bool
    moving: bool
    rest: np.ndarray   # (nv, 3) float32
    faces: np.ndarray  # (nf, 3) int32
    frames: list = field(default_factory=list)  # (nv, 3) per frame when moving


@dataclass
class Trajectory:
    times: np.ndarray
    objects: list


def read(path: str) -> Trajectory:
    with open(path, "rb") as f:
        data = f.read()
    if data[:8] != b"SILKTRJ1":
        raise ValueError(f"{path}: not a silk trajectory file")
    off = 8
    (nobj,) = struct.unpack_from("<I", data, off)
    off += 4
    objs = []
    for _ in range(nobj):
        kind, moving, nlen = struct.unpack_from("<III", data, off)
        off += 12
        name = data[off:off + nlen].decode("utf-8")
        off += nlen
        nv, nf = struct.unpack_from("<II", data, off)
        off += 8
        rest = np.frombuffer(data, np.float32, 3 * nv, off).reshape(nv, 3).copy()
        off += 12 * nv
        faces = np.frombuffer(data, np.int32, 3 * nf, off).reshape(nf, 3).copy()
        off += 12 * nf
        objs.append(TrajectoryObject(name, kind == 0, moving == 1, rest, faces))
    (nframes,) = struct.unpack_from("<I", data, off)
    off += 4
    times = np.zeros(nframes, np.float32)
    for k in range(nframes):
        (times[k],) = struct.unpack_from("<f", data, off)
        off += 4
        for o in objs:
            if o.moving:
                nv = o.rest.shape[0]
                o.frames.append(np.frombuffer(data, np.float32, 3 * nv, off).reshape(nv, 3).copy())
                off += 12 * nv
    if off != len(data):
        raise ValueError(f"{path}: {len(data) - off} trailing bytes")
    return Trajectory(times, objs)
