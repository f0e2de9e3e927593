"""Seeded particle scenes and the binary state snapshot format.

Replaces the reference's CPU scene sources (fluids/DoubleDamSource.cpp:5-41,
fluids/FixedCubeSource.cpp:6-36, fluids/ParticleSource.h:4-17), which fill a
cell-centred lattice (x outer, y, z inner; identities 0..N-1; zero
velocity) with rand() jitter seeded by srand(27).  Here the jitter comes from
SplitMix64 (counter form, so any slice can be generated independently) and is
uniform in [-jitter, +jitter] per axis; lattice points are cell-centred:
lo + (i + 1/2) * spacing.
"""
from __future__ import annotations

import dataclasses
import struct
from typing import List, Optional, Sequence, Tuple

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
SEED_BASE = 0x5EED0000


def splitmix64(seed: int, start: int, count: int) -> np.ndarray:
    """Outputs start .. start+count-1 of SplitMix64 seeded with `seed`
    (output k mixes seed + (k+1) * 0x9E3779B97F4A7C15)."""
    with np.errstate(over="ignore"):
        k = np.arange(start + 1, start + count + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + k * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform01(seed: int, start: int, count: int) -> np.ndarray:
    """Doubles in [0, 1) from the top 53 bits."""
    return (splitmix64(seed, start, count) >> np.uint64(11)).astype(np.float64) * (1.0 / 2**53)


def normal(seed: int, start: int, count: int) -> np.ndarray:
    """Standard normals by Box-Muller in FP64 (two uniforms per output)."""
    u = uniform01(seed, 2 * start, 2 * count).reshape(count, 2)
    r = np.sqrt(-2.0 * np.log1p(-u[:, 0]))
    return r * np.cos(2.0 * np.pi * u[:, 1])


@dataclasses.dataclass
class Scene:
    """Particle state in input order plus its box.  pos/vel: (N,3) float32."""

    pos: np.ndarray
    vel: np.ndarray
    iid: np.ndarray
    ulim: Tuple[float, float, float]
    llim: Tuple[float, float, float]
    name: str = ""

    @property
    def n(self) -> int:
        return int(self.pos.shape[0])

    def as_layout(self, layout: int) -> Tuple[np.ndarray, np.ndarray]:
        """pos, vel as (N, layout) float32 (w = 0 for layout 4)."""
        if layout == 3:
            return (np.ascontiguousarray(self.pos, np.float32),
                    np.ascontiguousarray(self.vel, np.float32))
        p = np.zeros((self.n, 4), np.float32)
        v = np.zeros((self.n, 4), np.float32)
        p[:, :3] = self.pos
        v[:, :3] = self.vel
        return p, v


def lattice(counts: Sequence[int], lo: Sequence[float], spacing: float, jitter: float,
            seed: int, stream_offset: int = 0) -> np.ndarray:
    """Jittered cell-centred lattice, x outer / y / z inner
    (fluids/DoubleDamSource.cpp:8-20), (N,3) float32."""
    nx, ny, nz = (int(c) for c in counts)
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    base = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1).astype(np.float64)
    n = base.shape[0]
    u = uniform01(seed, stream_offset, 3 * n).reshape(n, 3)
    x = np.asarray(lo, np.float64) + (base + 0.5) * spacing + (2.0 * u - 1.0) * jitter
    return x.astype(np.float32)


def double_dam(seed: int = SEED_BASE + 1, h: float = 0.1) -> Scene:
    """Two 32x16x32 (x,y,z) blocks of 16,384 particles, spacing 0.5h, jitter
    +-0.05h, in opposite floor corners of a 40h x 20h x 20h box (z up)."""
    s, jit = 0.5 * h, 0.05 * h
    a = lattice((32, 16, 32), (0.0, 0.0, 0.0), s, jit, seed, 0)
    b = lattice((32, 16, 32), (24 * h, 12 * h, 0.0), s, jit, seed, 3 * a.shape[0])
    pos = np.concatenate([a, b])
    return Scene(pos, np.zeros_like(pos), np.arange(pos.shape[0], dtype=np.uint32),
                 (40 * h, 20 * h, 20 * h), (0.0, 0.0, 0.0), "double_dam_32k")


def single_dam(seed: int = SEED_BASE + 2, h: float = 0.1,
               counts: Sequence[int] = (128, 64, 128),
               box_cells: Sequence[int] = (256, 128, 128)) -> Scene:
    """A 128x64x128 lattice (1,048,576 particles, spacing 0.5h, jitter
    +-0.05h) in the floor corner of a 256h x 128h x 128h box."""
    s, jit = 0.5 * h, 0.05 * h
    pos = lattice(counts, (0.0, 0.0, 0.0), s, jit, seed, 0)
    ul = tuple(float(np.float32(c * h)) for c in box_cells)
    return Scene(pos, np.zeros_like(pos), np.arange(pos.shape[0], dtype=np.uint32), ul,
                 (0.0, 0.0, 0.0), f"single_dam_{pos.shape[0]}")


def block(counts: Sequence[int], box: Sequence[float], seed: int, h: float = 0.1,
          lo: Sequence[float] = (0.0, 0.0, 0.0), vel_sigma: float = 0.0) -> Scene:
    """One jittered lattice block at `lo` in the box [0, box]; optional
    Gaussian velocities (sigma per component)."""
    pos = lattice(counts, lo, 0.5 * h, 0.05 * h, seed, 0)
    n = pos.shape[0]
    if vel_sigma > 0:
        vel = (vel_sigma * normal(seed ^ 0xA5A5, 0, 3 * n)).reshape(n, 3).astype(np.float32)
    else:
        vel = np.zeros_like(pos)
    return Scene(pos, vel, np.arange(n, dtype=np.uint32), tuple(float(b) for b in box),
                 (0.0, 0.0, 0.0), f"block_{n}")


def batch_scenes(nscenes: int = 64, seed: int = SEED_BASE + 3, h: float = 0.1,
                 nmin: int = 4096, nmax: int = 262144, fill=(0.2, 0.5)) -> List[Scene]:
    """Independent box scenes of uneven size.  Scene k: N_k =
    round(exp(U(ln nmin, ln nmax))) particles and fill fraction f_k = U(fill),
    both from SplitMix64(seed) outputs 2k and 2k+1; a cubic box [0, L]^3 with
    L = h * ceil(0.5 * cbrt(N_k / f_k)) (so N_k lattice sites at spacing 0.5h
    fill about f_k of it); the particles are the first N_k sites (x outer, z
    inner) of an m x m x ceil(N_k / m^2) lattice block, m = ceil(cbrt(N_k)),
    in the floor corner, spacing 0.5h, jitter +-0.05h, zero velocity."""
    u = uniform01(seed, 0, 2 * nscenes).reshape(nscenes, 2)
    out: List[Scene] = []
    for k in range(nscenes):
        n = int(round(np.exp(np.log(nmin) + u[k, 0] * (np.log(nmax) - np.log(nmin)))))
        f = fill[0] + u[k, 1] * (fill[1] - fill[0])
        cells = int(np.ceil(0.5 * np.cbrt(n / f)))
        m = int(np.ceil(np.cbrt(n)))
        nz = -(-n // (m * m))
        side = float(np.float32(cells * h))
        if m * 0.5 * h > side or nz * 0.5 * h > side:
            raise ValueError(f"scene {k}: block does not fit its box")
        pos = lattice((m, m, nz), (0.0, 0.0, 0.0), 0.5 * h, 0.05 * h,
                      seed ^ ((k + 1) << 32), 0)[:n]
        out.append(Scene(pos, np.zeros_like(pos), np.arange(n, dtype=np.uint32),
                         (side, side, side), (0.0, 0.0, 0.0), f"batch{k}_{n}"))
    return out


def differentiable(seed: int = SEED_BASE + 4, h: float = 0.1) -> Scene:
    """The differentiated workload: a 32x16x32 jittered lattice (16,384
    particles, spacing 0.5h, jitter +-0.05h) in the floor corner of a
    40h x 20h x 40h box, initial velocities N(0, 0.5^2) per component."""
    return dataclasses.replace(block((32, 16, 32), (40 * h, 20 * h, 40 * h), seed, h,
                                     vel_sigma=0.5), name="differentiable_16k")


def target_cloud(sc: Scene, seed: int = SEED_BASE + 0x40) -> np.ndarray:
    """A seeded target point per particle identity (N x 3, FP64): uniform in
    the lower half of the scene's box."""
    u = uniform01(seed, 0, 3 * sc.n).reshape(sc.n, 3)
    lo = np.asarray(sc.llim, np.float64)
    hi = np.asarray(sc.ulim, np.float64)
    hi = lo + (hi - lo) * np.array([1.0, 1.0, 0.5])
    return lo + u * (hi - lo)


def sweep_dam(seed: int = SEED_BASE + 5, h: float = 0.1) -> Scene:
    """The large sweeping-wall workload: a 256x64x256 jittered lattice
    (4,194,304 particles, spacing 0.5h, jitter +-0.05h) in the floor corner
    of a resting box 160h x 64h x 160h; the moving wall is sweep_of_sweep_dam()."""
    return dataclasses.replace(single_dam(seed, h, counts=(256, 64, 256),
                                          box_cells=(160, 64, 160)), name="sweep_dam_4m")


def sweep_of_sweep_dam(h: float = 0.1, frames_per_period: int = 120):
    """The +x wall of sweep_dam moves as ulim.x = 160h + 20h sin(2 pi f / 120)
    (period 2 s at dt = 1/60): (ulim0, llim0, a_ulim, a_llim, w, start)."""
    ul = (float(np.float32(160 * h)), float(np.float32(64 * h)), float(np.float32(160 * h)))
    return ul, (0.0, 0.0, 0.0), (20 * h, 0.0, 0.0), (0.0, 0.0, 0.0), \
        2.0 * np.pi / frames_per_period, 0


def concat(scenes: Sequence[Scene], layout: int):
    """The batched arrays of `scenes`: pos, vel (N_total, layout), iid, and
    the per-scene counts, boxes and offsets."""
    pv = [s.as_layout(layout) for s in scenes]
    pos = np.concatenate([p for p, _ in pv])
    vel = np.concatenate([v for _, v in pv])
    iid = np.concatenate([s.iid for s in scenes])
    counts = [s.n for s in scenes]
    offsets = np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.int64)
    return pos, vel, iid, counts, [s.ulim for s in scenes], [s.llim for s in scenes], offsets


# ---- binary snapshot -----------------------------------------------------
# Little-endian: magic b"PBFS", u32 version=1, u32 n, u32 stride (3|4),
# f32 ulim[3], f32 llim[3], then pos (n*stride f32), vel (n*stride f32),
# iid (n u32).
_MAGIC = b"PBFS"
_HDR = struct.Struct("<4sIII3f3f")


def save_snapshot(path: str, pos: np.ndarray, vel: np.ndarray, iid: np.ndarray,
                  ulim: Sequence[float], llim: Sequence[float]) -> None:
    pos = np.ascontiguousarray(pos, dtype="<f4")
    vel = np.ascontiguousarray(vel, dtype="<f4")
    iid = np.ascontiguousarray(iid, dtype="<u4")
    n, stride = pos.shape
    if vel.shape != pos.shape or iid.shape != (n,) or stride not in (3, 4):
        raise ValueError("pos/vel must be (n,3|4) and iid (n,)")
    with open(path, "wb") as f:
        f.write(_HDR.pack(_MAGIC, 1, n, stride, *map(float, ulim), *map(float, llim)))
        f.write(pos.tobytes())
        f.write(vel.tobytes())
        f.write(iid.tobytes())


def load_snapshot(path: str):
    """Returns (pos, vel, iid, ulim, llim)."""
    with open(path, "rb") as f:
        hdr = f.read(_HDR.size)
        if len(hdr) != _HDR.size:
            raise ValueError("truncated snapshot header")
        magic, ver, n, stride, *lims = _HDR.unpack(hdr)
        if magic != _MAGIC or ver != 1 or stride not in (3, 4):
            raise ValueError("not a PBFS v1 snapshot")
        body = f.read()
    need = 4 * (2 * n * stride + n)
    if len(body) != need:
        raise ValueError(f"snapshot body has {len(body)} bytes, expected {need}")
    a = np.frombuffer(body, dtype="<f4", count=2 * n * stride)
    pos = a[: n * stride].reshape(n, stride).copy()
    vel = a[n * stride:].reshape(n, stride).copy()
    iid = np.frombuffer(body, dtype="<u4", offset=8 * n * stride, count=n).copy()
    return pos, vel, iid, tuple(lims[:3]), tuple(lims[3:])
