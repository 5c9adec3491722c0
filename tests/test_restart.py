"""Restart files (haccsr.restart): the serial layout of
Particles::writeRestart and the block layout of RestartIO, byte for byte
against files assembled field by field here, and round trips.  The
reference ships no restart file, so the layouts are pinned to its write
calls (Particles.cxx:1070-1080, RestartIO.cxx:396-411, 196-208)."""
import struct

import numpy as np
import pytest

from haccsr import restart as R


def parts(n, seed):
    rng = np.random.default_rng(seed)
    f = {k: rng.normal(size=n).astype(np.float32) for k in ("x", "y", "z", "vx", "vy", "vz", "phi")}
    return R.RestartParticles(**f, id=rng.integers(0, 2 ** 40, n).astype(np.int64),
                              mask=rng.integers(0, 65535, n).astype(np.uint16),
                              mass=np.ones(n, np.float32))


def serial_bytes(p):
    b = struct.pack("<i", p.n)
    for k in ("x", "vx", "y", "vy", "z", "vz", "phi"):
        b += struct.pack(f"<{p.n}f", *getattr(p, k))
    b += struct.pack(f"<{p.n}q", *p.id) + struct.pack(f"<{p.n}H", *p.mask)
    return b


def block_bytes(ps):
    out, starts = b"", []
    for p in ps:
        starts.append(len(out))
        out += struct.pack("<i", p.n)
        for k in ("x", "y", "z", "vx", "vy", "vz", "phi"):
            out += struct.pack(f"<{p.n}f", *getattr(p, k))
        out += struct.pack(f"<{p.n}q", *p.id) + struct.pack(f"<{p.n}H", *p.mask)
    return out + struct.pack(f"<{len(ps) + 1}i", *starts, len(ps))


def same(a, b):
    for k in ("x", "y", "z", "vx", "vy", "vz", "phi", "id", "mask"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k


@pytest.mark.parametrize("n", [0, 1, 1000])
def test_serial_layout_and_round_trip(tmp_path, n):
    p = parts(n, n)
    f = tmp_path / "r.bin"
    R.write_restart(f, p)
    assert f.read_bytes() == serial_bytes(p)
    q = R.read_restart(f)
    same(p, q)
    assert np.all(q.mass == 1)


def test_block_layout_footer_and_ranks(tmp_path):
    ps = [parts(n, 10 + i) for i, n in enumerate((300, 0, 41, 1000))]
    f = tmp_path / "b.bin"
    R.write_restart(f, ps, mpiio=True)
    assert f.read_bytes() == block_bytes(ps)
    starts = R.read_footer(f)
    assert list(starts) == [0, 4 + 38 * 300, 2 * 4 + 38 * 300, 3 * 4 + 38 * 341]
    for r, p in enumerate(ps):
        same(p, R.read_restart(f, mpiio=True, rank=r))
    allp = R.read_restart(f, mpiio=True)
    assert allp.n == 1341


def test_truncated_serial_file_is_an_error(tmp_path):
    f = tmp_path / "t.bin"
    R.write_restart(f, parts(10, 3))
    f.write_bytes(f.read_bytes() + b"\0\0")
    with pytest.raises(ValueError):
        R.read_restart(f)
