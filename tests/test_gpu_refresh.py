"""GPU parity of the one-rank overload refresh (hsr_refresh) with the oracle:
copyAliveIntoVectors, ParticleExchange with every neighbour this rank,
copyAllFromVectors and resortParticles (MC3Extras.cxx:330-338, 660-700).
Every array, the refreshed count, m_Np_last and the alive count bit-exact;
host and device arrays; the capacity error; the C++ drop-in; and a
clustered load, where a second refresh of unmoved particles must give
the same particles (positions, ids, masks) again.  The 256^3 load is timed
by bench.py's refresh leg."""
import os
import subprocess
import sys

import numpy as np
import pytest

from haccsr import _lib, ics
from haccsr.particles import Particles
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ARR = ("x", "y", "z", "vx", "vy", "vz", "mass", "phi", "id", "mask")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def bits_equal(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.dtype.kind == "f":
        return np.array_equal(a.view(np.uint32), b.view(np.uint32))
    return np.array_equal(a, b)


def state(loc, seed):
    rng = np.random.default_rng(seed)
    n = loc.shape[0]
    return dict(x=loc[:, 0].copy(), y=loc[:, 1].copy(), z=loc[:, 2].copy(),
                vx=rng.normal(size=n).astype(np.float32),
                vy=rng.normal(size=n).astype(np.float32),
                vz=rng.normal(size=n).astype(np.float32),
                mass=rng.uniform(0.5, 2, n).astype(np.float32),
                phi=rng.normal(size=n).astype(np.float32),
                id=rng.permutation(n).astype(np.int64) * 3 + 1,
                mask=rng.integers(0, 65535, n).astype(np.uint16))


def padded(st, cap):
    out = {}
    for k, a in st.items():
        b = np.zeros(cap, a.dtype)
        b[:a.shape[0]] = a
        out[k] = b
    return out


def uniform_local(n, ngt, seed):
    """Local grid coordinates over the whole overloaded grid and a little
    beyond it (old copies, particles that left the box)."""
    rng = np.random.default_rng(seed)
    return rng.uniform(-1.0, ngt + 1.0, (n, 3)).astype(np.float32)


def to_dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def check(st, count, rL, oL, ng, anow, cap, device=False, ctx=None):
    ref, rcount, rinside, ralive = O.refresh(padded(st, cap), count, rL, oL, ng, anow)
    own = ctx is None
    ctx = ctx or _lib.Context()
    dom = _lib.domain(rL, oL, ng)
    buf = padded(st, cap)
    if device:
        import torch
        d = {k: to_dev(v) for k, v in buf.items()}
        n2, inside = ctx.refresh(dom, anow, count, *(d[k] for k in ARR), device_ptrs=True)
        torch.cuda.synchronize()
        got = {k: d[k].cpu().numpy() for k in ARR}
    else:
        n2, inside = ctx.refresh(dom, anow, count, *(buf[k] for k in ARR))
        got = buf
    stt = ctx.stats()
    assert (n2, inside) == (rcount, rinside), ((n2, inside), (rcount, rinside))
    assert stt.refresh_alive == ralive and stt.refresh_dead == rcount - ralive
    for k in ARR:
        assert bits_equal(got[k][:n2], ref[k][:n2]), k
    if own:
        ctx.close()
    return got, n2, inside, ralive, stt.refresh_ms


@pytest.mark.parametrize("rL,oL,ng,anow,n,device", [
    (64.0, 3.2, 64, 1.0, 300000, False),
    (100.0, 8.0, 64, 0.37, 200000, True),
    (1000.0, 10.0, 512, 0.5, 50000, False),
])
def test_refresh_matches_oracle(rL, oL, ng, anow, n, device):
    say(f"case refresh rL={rL} oL={oL} ng={ng} anow={anow} n={n} device={device}")
    d = _lib.domain(rL, oL, ng)
    st = state(uniform_local(n, d.ng_local_total, 3), 4)
    _, cnt, inside, alive, ms = check(st, n, rL, oL, ng, anow, 2 * n, device=device)
    say(f"  alive {alive} copies {cnt - alive} inside {inside} ({ms:.3f} ms)")
    assert alive > 0 and cnt > alive


def test_refresh_boundaries_empty_and_capacity():
    rL, oL = 32.0, 3.0
    d = _lib.domain(rL, oL, 32)
    og = np.float32(d.ng_overload)
    vals = np.array([-0.5, 0.0, 0.25, oL, oL + 0.125, rL / 2, rL - oL - 0.125, rL - oL,
                     rL - 0.25, rL, rL + 0.5], np.float32)
    g = np.stack(np.meshgrid(vals, vals, vals, indexing="ij"), -1).reshape(-1, 3)
    st = state((g + og).astype(np.float32), 9)
    n = g.shape[0]
    say("case refresh boundary values")
    ctx = _lib.Context()
    got, cnt, inside, alive, _ = check(st, n, rL, oL, 32, 1.0, 8 * n, ctx=ctx)
    say("case refresh empty")
    _, cnt0, inside0, alive0, _ = check(st, 0, rL, oL, 32, 1.0, n, ctx=ctx)
    assert (cnt0, inside0, alive0) == (0, 0, 0)
    say("case refresh capacity too small: error, arrays untouched")
    buf = padded(st, cnt - 1)
    keep = {k: v.copy() for k, v in buf.items()}
    with pytest.raises(_lib.CapacityError) as e:
        ctx.refresh(d, 1.0, n, *(buf[k] for k in ARR))
    assert e.value.required == cnt
    for k in ARR:
        assert bits_equal(buf[k], keep[k]), k
    ctx.close()


def test_refresh_clustered_128_idempotent():
    """128^3 clustered particles (no copies yet): parity with the oracle,
    then a second refresh of the unmoved result gives the same positions,
    ids and masks in the same order (velocities go through the unit factors
    once more)."""
    import torch
    say("case refresh 128^3 clustered, device arrays")
    rL, oL, ng = 128.0, 3.2, 128
    d = _lib.domain(rL, oL, ng)
    x, y, z = ics.clustered(128, 128.0, 62, 1004)
    og = np.float32(d.ng_overload)
    loc = np.stack([x, y, z], 1).astype(np.float32) + og
    n = loc.shape[0]
    st = state(loc, 5)
    cap = n + n // 4
    ctx = _lib.Context()
    got, cnt, inside, alive, ms = check(st, n, rL, oL, ng, 1.0, cap, device=True, ctx=ctx)
    say(f"  alive {alive} copies {cnt - alive} inside {inside}: {ms:.2f} ms")
    assert 0.999 * n < alive <= n and inside == cnt
    dev = {k: to_dev(got[k]) for k in ARR}
    n2, inside2 = ctx.refresh(d, 1.0, cnt, *(dev[k] for k in ARR), device_ptrs=True)
    torch.cuda.synchronize()
    assert (n2, inside2) == (cnt, inside)
    for k in ("x", "y", "z", "id", "mask", "mass", "phi"):
        assert bits_equal(dev[k][:n2].cpu().numpy(), got[k][:cnt]), k
    say(f"  second refresh identical: {ctx.stats().refresh_ms:.2f} ms")
    ctx.close()


def test_particles_mirror_refresh_grows_arrays():
    """Particles.refreshParticles: new arrays of the refreshed length, the
    oracle's result, with the capacity grown when the first guess is short."""
    rL, oL, ng = 48.0, 6.0, 48
    d = _lib.domain(rL, oL, ng)
    st = state(uniform_local(60000, d.ng_local_total, 6), 7)
    ref, rc, ri, ra = O.refresh(padded(st, 4 * 60000), 60000, rL, oL, ng, 0.9)
    say("case Particles.refreshParticles (capacity grown)")
    P = Particles([d.ng_local_total] * 3, alpha=1.0, gpscal=1.0, edge=2.0)
    a = {k: v.copy() for k, v in st.items()}
    P.set_particles(a["x"], a["y"], a["z"], a["vx"], a["vy"], a["vz"], a["mass"], a["phi"],
                    a["id"], a["mask"])
    cnt = P.refreshParticles(rL, oL, ng, 0.9, capacity=1000)
    assert cnt == rc and P.Np_last == ri and P.Np_local_total == rc
    for k in ARR:
        assert bits_equal(P.arrays[k], ref[k][:rc]), k
    P.close()


def test_cxx_dropin_refresh(tmp_path):
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "hacc-coral_amd", "bin", "test_subcycle")
    n, g = 40000, 32
    out = tmp_path / "refresh.bin"
    say("case C++ ParticleSubCycle::refreshParticles")
    r = subprocess.run([exe, str(n), str(g), "64", "91", "2", str(out)], check=True,
                       timeout=120, capture_output=True, text=True)
    fields = dict(kv.split("=") for kv in r.stdout.split() if "=" in kv)
    cap = 3 * n
    raw = np.fromfile(out, dtype=np.uint8)
    assert raw.shape[0] == 42 * (n + cap)
    parts = []
    for part, m in ((raw[:42 * n], n), (raw[42 * n:], cap)):
        f = part[: 32 * m].view(np.float32).reshape(8, m)
        dd = dict(zip(("x", "y", "z", "vx", "vy", "vz", "phi", "mass"), f))
        dd["id"] = part[32 * m: 40 * m].view(np.int64)
        dd["mask"] = part[40 * m: 42 * m].view(np.uint16)
        parts.append(dd)
    init, fin = parts
    st = {k: v[:n].copy() for k, v in init.items()}
    ref, rc, ri, ra = O.refresh(padded(st, cap), n, float(g), 2.5, g, 0.8)
    assert int(fields["count"]) == rc and int(fields["Np_last"]) == ri
    for k in ARR:
        assert bits_equal(fin[k][:rc], ref[k][:rc]), k


def test_refresh_twice_reuses_buffers():
    """Two refreshes on one context, the second with fewer particles: the
    buffers sized by the first serve the second, results as the oracle's."""
    rL, oL, ng = 64.0, 3.2, 64
    d = _lib.domain(rL, oL, ng)
    ctx = _lib.Context()
    for n, seed in ((250000, 41), (30000, 42)):
        say(f"case refresh twice: n={n}")
        st = state(uniform_local(n, d.ng_local_total, seed), seed + 1)
        check(st, n, rL, oL, ng, 1.0, 2 * n, device=(n < 100000), ctx=ctx)
    ctx.close()
