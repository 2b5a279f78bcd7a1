"""NBS2 trajectory files (the reference's output format, src/Main.cpp:13-49,
read by load_sim.py:47-131).

Layout, little-endian:
  header (28 bytes): "NBS2", u32 n_bodies, u32 total_steps, u32 output_every,
                     u32 num_frames, u32 completed_frames (updated live),
                     f32 dt
  f32 mass[n], f32 radius[n]
  num_frames frames of n records [x, y, vx, vy] (f32), in id order; frame 0
  is the initial state, frame k the state after step k * output_every.

`read` trusts completed_frames (a killed writer leaves a readable, shorter
file) and falls back to the frames actually present.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass

import numpy as np

MAGIC = b"NBS2"
HEADER = struct.Struct("<4sIIIIIf")


@dataclass
class Trajectory:
    n_bodies: int
    total_steps: int
    output_every: int
    num_frames: int
    completed_frames: int
    dt: float
    mass: np.ndarray     # (n,)
    radius: np.ndarray   # (n,)
    pos: np.ndarray      # (frames, n, 2)
    vel: np.ndarray      # (frames, n, 2)


def read(path: str) -> Trajectory:
    with open(path, "rb") as f:
        raw = f.read()
    if len(raw) < HEADER.size:
        raise ValueError("file shorter than the NBS2 header")
    magic, n, steps, every, frames, done, dt = HEADER.unpack_from(raw, 0)
    if magic != MAGIC:
        raise ValueError(f"bad magic {magic!r}")
    off = HEADER.size
    mass = np.frombuffer(raw, "<f4", n, off)
    radius = np.frombuffer(raw, "<f4", n, off + 4 * n)
    off += 8 * n
    per = 16 * n
    present = (len(raw) - off) // per if per else 0
    k = min(done, present) if done <= frames else present
    data = np.frombuffer(raw, "<f4", k * n * 4, off).reshape(k, n, 4)
    return Trajectory(n, steps, every, frames, done, dt, mass.copy(), radius.copy(),
                      data[:, :, 0:2].copy(), data[:, :, 2:4].copy())


def write_header(f, n: int, total_steps: int, output_every: int, dt: float) -> int:
    """Writes the header with completed_frames = 0; returns num_frames."""
    frames = 1 + total_steps // output_every
    f.write(HEADER.pack(MAGIC, n, total_steps, output_every, frames, 0, dt))
    return frames
