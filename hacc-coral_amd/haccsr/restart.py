"""Particle restart files of the reference, read into and written from the
arrays the device path takes.

Two layouts (Particles::writeRestart / readRestart, Particles.cxx:1059-1123):

* serial (``mpiio=False``): ``int32 Np``, then ``x, vx, y, vy, z, vz, phi``
  as float32[Np] each, ``id`` as int64[Np] and ``mask`` as uint16[Np];
* block file (``mpiio=True``, RestartIO.cxx:69-418): one block per rank,
  each ``int32 Np`` then ``x, y, z, vx, vy, vz, phi`` float32[Np], ``id``
  int64[Np], ``mask`` uint16[Np], blocks back to back in rank order,
  followed by a footer of the blocks' byte offsets (int32 per block) and
  the block count (int32).

Both are native little-endian, as the reference writes them on x86.  Masses
are not stored; readRestart sets them to 1.
"""
from __future__ import annotations

import dataclasses

import numpy as np

SERIAL_ORDER = ("x", "vx", "y", "vy", "z", "vz", "phi")
BLOCK_ORDER = ("x", "y", "z", "vx", "vy", "vz", "phi")


@dataclasses.dataclass
class RestartParticles:
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    vx: np.ndarray
    vy: np.ndarray
    vz: np.ndarray
    phi: np.ndarray
    id: np.ndarray
    mask: np.ndarray
    mass: np.ndarray

    @property
    def n(self) -> int:
        return int(self.x.shape[0])


def _make(fields: dict) -> RestartParticles:
    n = fields["x"].shape[0]
    return RestartParticles(**{k: np.ascontiguousarray(v) for k, v in fields.items()},
                            mass=np.ones(n, np.float32))


def _block_bytes(n: int) -> int:
    return 4 + n * (7 * 4 + 8 + 2)


def write_restart(path, parts, mpiio: bool = False):
    """Write one particle set (serial layout) or a list of per-rank sets
    (block layout)."""
    with open(path, "wb") as f:
        if not mpiio:
            p = parts
            f.write(np.int32(p.n).tobytes())
            for k in SERIAL_ORDER:
                f.write(np.ascontiguousarray(getattr(p, k), np.float32).tobytes())
            f.write(np.ascontiguousarray(p.id, np.int64).tobytes())
            f.write(np.ascontiguousarray(p.mask, np.uint16).tobytes())
            return
        blocks = list(parts)
        starts = []
        off = 0
        for p in blocks:
            starts.append(off)
            f.write(np.int32(p.n).tobytes())
            for k in BLOCK_ORDER:
                f.write(np.ascontiguousarray(getattr(p, k), np.float32).tobytes())
            f.write(np.ascontiguousarray(p.id, np.int64).tobytes())
            f.write(np.ascontiguousarray(p.mask, np.uint16).tobytes())
            off += _block_bytes(p.n)
        f.write(np.asarray(starts + [len(blocks)], np.int32).tobytes())


def read_footer(path) -> np.ndarray:
    """Block start offsets of a block-layout file (RestartIO::ReadFooter)."""
    raw = np.memmap(path, dtype=np.uint8, mode="r")
    nb = int(raw[-4:].view(np.int32)[0])
    if nb < 0 or 4 * (nb + 1) > raw.shape[0]:
        raise ValueError(f"{path}: not a block restart file (footer count {nb})")
    return np.array(raw[-4 * (nb + 1):-4].view(np.int32))


def read_restart(path, mpiio: bool = False, rank: int | None = None) -> RestartParticles:
    """Read a serial file, or one rank's block (all blocks concatenated when
    ``rank`` is None) of a block file."""
    raw = np.memmap(path, dtype=np.uint8, mode="r")
    if not mpiio:
        n = int(raw[:4].view(np.int32)[0])
        o = 4
        out = {}
        for k in SERIAL_ORDER:
            out[k] = np.array(raw[o:o + 4 * n].view(np.float32))
            o += 4 * n
        out["id"] = np.array(raw[o:o + 8 * n].view(np.int64))
        o += 8 * n
        out["mask"] = np.array(raw[o:o + 2 * n].view(np.uint16))
        o += 2 * n
        if o != raw.shape[0]:
            raise ValueError(f"{path}: {raw.shape[0] - o} trailing bytes after {n} particles")
        return _make(out)
    starts = read_footer(path)
    ranks = range(starts.shape[0]) if rank is None else [rank]
    parts = []
    for r in ranks:
        o = int(starts[r])
        n = int(raw[o:o + 4].view(np.int32)[0])
        o += 4
        out = {}
        for k in BLOCK_ORDER:
            out[k] = np.array(raw[o:o + 4 * n].view(np.float32))
            o += 4 * n
        out["id"] = np.array(raw[o:o + 8 * n].view(np.int64))
        o += 8 * n
        out["mask"] = np.array(raw[o:o + 2 * n].view(np.uint16))
        parts.append(out)
    return _make({k: np.concatenate([p[k] for p in parts]) for k in parts[0]})
