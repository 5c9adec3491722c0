"""The sub-cycle on the device (hsr_drift, hsr_resort, hsr_subcycle): drift
and resort bit-exact against the oracle, the one-call device-resident cycle
bit-exact against the same cycle run as single steps, and each sub-step of
that cycle against the oracle's sub-step from the same state (drift, resort,
tree and lists exact; kick and the drift after it to the force tolerance)."""
import numpy as np
import pytest

from haccsr import Context
from haccsr.particles import Particles, TimeStepperValues
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TS = TimeStepperValues(pp=0.5, tau=0.02, tau2=0.02, adot=1.3, fscal=1.1)
ARR = ("x", "y", "z", "vx", "vy", "vz", "mass", "phi", "id", "mask")


def say(msg):
    print(msg, flush=True)


def bits_equal(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.dtype == np.float32:
        return np.array_equal(a.view(np.uint32), b.view(np.uint32))
    return np.array_equal(a, b)


def to_dev(v):
    import torch
    a = v.copy()
    return torch.from_numpy(a.view(np.int16) if a.dtype == np.uint16 else a).cuda()


def to_host(t, like):
    a = t.cpu().numpy()
    return a.view(np.uint16) if like.dtype == np.uint16 else a


def particles(n, ng, seed, spread=1.5, vscale=4.0):
    """Uniform particles plus clumps, a shell of them outside the grid."""
    rng = np.random.default_rng(seed)
    g = np.asarray(ng, np.float64)
    nu = n // 2
    p = rng.uniform(0, 1, (nu, 3)) * (g + 2 * spread) - spread
    centres = rng.uniform(0.15, 0.85, (12, 3)) * g
    k = rng.integers(0, 12, n - nu)
    c = centres[k] + rng.normal(0, 0.6, (n - nu, 3))
    pos = np.concatenate([p, c]).astype(np.float32)
    pos = pos[rng.permutation(n)]
    st = {"x": pos[:, 0].copy(), "y": pos[:, 1].copy(), "z": pos[:, 2].copy()}
    for a in ("vx", "vy", "vz"):
        st[a] = rng.normal(0, vscale, n).astype(np.float32)
    st["mass"] = rng.uniform(0.5, 2.0, n).astype(np.float32)
    st["phi"] = rng.normal(0, 1, n).astype(np.float32)
    st["id"] = (np.arange(n, dtype=np.int64) * 7 + 3)
    st["mask"] = rng.integers(0, 65535, n).astype(np.uint16)
    return st


def test_drift_bitexact():
    ctx = Context()
    for n, seed in ((0, 1), (1, 2), (7, 3), (100003, 4)):
        say(f"case drift n={n}")
        st = particles(max(n, 1), [32, 32, 32], seed)
        st = {a: v[:n].copy() for a, v in st.items()}
        for pp, tau, adot, alpha in [(0.5, 0.004, 1.3, 1.0), (0.21, 0.01, 3.7, 1.5)]:
            k = O.map1_factor(pp, tau, adot, alpha)
            want = O.drift(st["x"], st["y"], st["z"], st["vx"], st["vy"], st["vz"], k)
            X, Y, Z = st["x"].copy(), st["y"].copy(), st["z"].copy()
            ctx.drift(X, Y, Z, st["vx"], st["vy"], st["vz"], pp, tau, adot, alpha)
            for g, w in zip((X, Y, Z), want):
                assert bits_equal(g, w)
    ctx.close()


def test_drift_device_tensors_unaligned():
    """Device arrays at an offset that is not 16-byte aligned take the scalar
    path; the result is the same."""
    import torch
    say("case drift device tensors, offset 1")
    n = 50001
    st = particles(n + 1, [32, 32, 32], 9)
    k = O.map1_factor(0.5, 0.004, 1.3, 1.0)
    want = O.drift(st["x"][1:], st["y"][1:], st["z"][1:], st["vx"][1:], st["vy"][1:],
                   st["vz"][1:], k)
    d = {a: torch.from_numpy(st[a]).cuda() for a in ("x", "y", "z", "vx", "vy", "vz")}
    ctx = Context()
    ctx.drift(d["x"][1:], d["y"][1:], d["z"][1:], d["vx"][1:], d["vy"][1:], d["vz"][1:],
              0.5, 0.004, 1.3, 1.0, device_ptrs=True)
    torch.cuda.synchronize()
    for a, w in zip(("x", "y", "z"), want):
        assert bits_equal(d[a][1:].cpu().numpy(), w)
    ctx.close()


@pytest.mark.parametrize("n,ng,seed", [(0, (8, 8, 8), 1), (1, (8, 8, 8), 2),
                                       (20000, (12, 9, 7), 3), (400000, (40, 40, 40), 4)])
def test_resort_bitexact(n, ng, seed):
    say(f"case resort n={n} ng={ng}")
    st = particles(max(n, 1), ng, seed)
    st = {a: v[:n].copy() for a, v in st.items()}
    if n > 100:
        # cell faces, -0.0, the far face and duplicates of one cell
        st["x"][:6] = np.array([0.0, -0.0, ng[0], np.nextafter(ng[0], 0), -1e-7, 3.0], np.float32)
        st["y"][50:90] = 2.5
        st["z"][50:90] = 1.25
        st["x"][50:90] = 4.75
    want, inside = O.resort(st, ng)
    got = {a: v.copy() for a, v in st.items()}
    ctx = Context()
    g_inside = ctx.resort(ng, got["x"], got["y"], got["z"], got["vx"], got["vy"], got["vz"],
                          got["mass"], got["phi"], got["id"], got["mask"])
    ctx.close()
    assert g_inside == inside
    for a in ARR:
        assert bits_equal(got[a], want[a]), a


def test_resort_partial_arrays_and_device():
    """Only positions and ids present, on the device."""
    import torch
    say("case resort device, positions + ids")
    ng = (16, 16, 16)
    st = particles(123457, ng, 5)
    sub = {a: st[a] for a in ("x", "y", "z", "id")}
    want, inside = O.resort(sub, ng)
    d = {a: to_dev(v) for a, v in sub.items()}
    ctx = Context()
    g_inside = ctx.resort(ng, d["x"], d["y"], d["z"], ids=d["id"], device_ptrs=True)
    torch.cuda.synchronize()
    ctx.close()
    assert g_inside == inside
    for a in sub:
        assert bits_equal(d[a].cpu().numpy(), want[a]), a


def test_resort_rejects_bad_grid():
    from haccsr import HsrError
    ctx = Context()
    x = np.zeros(4, np.float32)
    with pytest.raises(HsrError):
        ctx.resort([0, 4, 4], x, x.copy(), x.copy())
    with pytest.raises(HsrError):
        ctx.resort([2048, 2048, 1024], x, x.copy(), x.copy())
    ctx.close()


SP = dict(alpha=1.0, gpscal=1.0, edge=2.0, nsub=5)
TREE = dict(fsrrmax=3.12, rsm=0.1, oa=0.1)


def make(ng, nd, **kw):
    return Particles(ng, rcb_tree_ppn=nd, **{**SP, **kw}, **TREE)


def test_subcycle_matches_single_steps():
    """hsr_subcycle (particles on the device for the whole cycle) equals the
    same cycle as map1 / map2 / map1 calls, host and device arrays."""
    import torch
    ng, nd = (24, 24, 24), 64
    st = particles(40000, ng, 11)
    say("case subcycle one call vs single steps (host arrays)")
    a = {k: v.copy() for k, v in st.items()}
    P1 = make(ng, nd)
    P1.set_particles(a["x"], a["y"], a["z"], a["vx"], a["vy"], a["vz"], a["mass"], a["phi"],
                     a["id"], a["mask"])
    P1.subCycle(TS)
    b = {k: v.copy() for k, v in st.items()}
    P2 = make(ng, nd)
    P2.set_particles(b["x"], b["y"], b["z"], b["vx"], b["vy"], b["vz"], b["mass"], b["phi"],
                     b["id"], b["mask"])
    P2.subCycle_steps(TS)
    assert P1.Np_last == P2.Np_last > 0
    for k in ARR:
        assert bits_equal(a[k], b[k]), k
    v0 = st["vx"][np.searchsorted(st["id"], a["id"])]
    assert np.abs(a["vx"] - v0).max() > 0  # the kicks did act
    say("case subcycle one call, device tensors")
    d = {k: to_dev(v) for k, v in st.items()}
    P3 = make(ng, nd)
    P3.set_particles(d["x"], d["y"], d["z"], d["vx"], d["vy"], d["vz"], d["mass"], d["phi"],
                     d["id"], d["mask"])
    P3.subCycle(TS)
    torch.cuda.synchronize()
    assert P3.Np_last == P1.Np_last
    for k in ARR:
        assert bits_equal(to_host(d[k], a[k]), a[k]), k
    for P in (P1, P2, P3):
        P.close()


def test_subcycle_skips():
    """skip_kick leaves velocities alone (drifts only); skip_stream leaves
    positions to the resort and tree permutations."""
    ng, nd = (16, 16, 16), 64
    st = particles(20000, ng, 13)
    say("case subcycle skip_kick / skip_stream")
    a = {k: v.copy() for k, v in st.items()}
    P = make(ng, nd, skip_kick=True)
    P.set_particles(a["x"], a["y"], a["z"], a["vx"], a["vy"], a["vz"], a["mass"], a["phi"],
                    a["id"], a["mask"])
    P.subCycle(TS)
    assert P.Np_last == -1
    k = O.map1_factor(TS.pp, np.float32(TS.tau2 / 5), TS.adot, 1.0)
    x, y, z = st["x"], st["y"], st["z"]
    for _ in range(10):
        x, y, z = O.drift(x, y, z, st["vx"], st["vy"], st["vz"], k)
    assert bits_equal(a["x"], x) and bits_equal(a["vz"], st["vz"])
    P.close()
    b = {k: v.copy() for k, v in st.items()}
    P = make(ng, nd, skip_stream=True)
    P.set_particles(b["x"], b["y"], b["z"], b["vx"], b["vy"], b["vz"], b["mass"], b["phi"],
                    b["id"], b["mask"])
    P.subCycle(TS)
    order = np.searchsorted(st["id"], b["id"])
    assert bits_equal(b["x"], st["x"][order]) and bits_equal(b["z"], st["z"][order])
    assert np.all(b["mass"] == 1.0)
    P.close()


def test_step_factors_bitexact():
    """The drift factor and kick coefficient the library derives on the host
    from the TimeStepper values equal the oracle's bit for bit."""
    ng, nd = (12, 12, 12), 64
    st = particles(3000, ng, 19)
    rng = np.random.default_rng(4)
    for trial in range(6):
        ts = TimeStepperValues(pp=rng.uniform(0.05, 2.0), tau=rng.uniform(1e-3, 0.1),
                               tau2=rng.uniform(1e-3, 0.1), adot=rng.uniform(0.1, 4.0),
                               fscal=rng.uniform(0.2, 3.0))
        nsub = int(rng.integers(1, 7))
        alpha, gpscal = float(rng.choice([0.5, 1.0, 1.5, 2.0])), float(rng.uniform(0.5, 2.5))
        say(f"case step factors trial {trial} nsub {nsub}")
        a = {k: v.copy() for k, v in st.items()}
        P = make(ng, nd, nsub=nsub, alpha=alpha, gpscal=gpscal)
        P.set_particles(a["x"], a["y"], a["z"], a["vx"], a["vy"], a["vz"], a["mass"], a["phi"],
                        a["id"], a["mask"])
        P.subCycle(ts)
        s = P.ctx.stats()
        sf = 1.0 / nsub
        k = O.map1_factor(np.float32(ts.pp), np.float32(sf * ts.tau2), np.float32(ts.adot), alpha)
        c = O.map2_coeff(np.float32(gpscal), ts.fscal, ts.tau, sf)
        assert np.float32(s.drift_factor) == np.float32(k)
        assert np.float32(s.kick_coeff) == np.float32(c)
        P.map2(ts, 0.3)
        assert np.float32(P.ctx.stats().kick_coeff) == np.float32(
            O.map2_coeff(np.float32(gpscal), ts.fscal, ts.tau, 0.3))
        P.close()


def test_subcycle_steps_against_oracle():
    """Each sub-step of the device cycle against the oracle's sub-step from
    the same starting state: the resort and the tree permutation are exact
    (ids in the same order, m_Np_last equal), velocities agree to the force
    tolerance (relative RMS of the kick 1e-5) and positions to the drift of
    that velocity difference."""
    ng, nd = (24, 24, 24), 64
    st = particles(40000, ng, 17)
    a = {k: v.copy() for k, v in st.items()}
    P = make(ng, nd)
    P.set_particles(a["x"], a["y"], a["z"], a["vx"], a["vy"], a["vz"], a["mass"], a["phi"],
                    a["id"], a["mask"])
    sp = dict(SP, ng=list(ng), pp=TS.pp, tau=TS.tau, tau2=TS.tau2, adot=TS.adot, fscal=TS.fscal)
    before = {"s": {k: v.copy() for k, v in a.items()}}

    def check(step):
        say(f"case subcycle step {step} vs oracle")
        want, inside, tree, lists = O.substep(before["s"], sp, dict(TREE, nd=nd))
        assert P.Np_last == inside
        assert bits_equal(a["id"], want["id"])
        assert bits_equal(a["mass"], want["mass"]) and bits_equal(a["mask"], want["mask"])
        V = np.stack([a[k] for k in ("vx", "vy", "vz")], 1).astype(np.float64)
        W = np.stack([want[k] for k in ("vx", "vy", "vz")], 1).astype(np.float64)
        # the kick each side applied, against the same pre-kick velocities
        pre = {k: before["s"][k] for k in ("id", "vx", "vy", "vz")}
        idx = np.argsort(pre["id"])
        src = idx[np.searchsorted(pre["id"], a["id"], sorter=idx)]
        V0 = np.stack([pre[k][src] for k in ("vx", "vy", "vz")], 1).astype(np.float64)
        kick_w = W - V0
        err = np.sqrt(np.sum((V - W) ** 2) / np.sum(kick_w ** 2))
        say(f"  inside {inside}, pairs {lists.pairs(tree):.3e}, kick rel rms {err:.2e}")
        assert np.sum(kick_w ** 2) > 0 and err < 1e-5
        kf = O.map1_factor(TS.pp, np.float32(TS.tau2 / 5), TS.adot, 1.0)
        for c, v in (("x", 0), ("y", 1), ("z", 2)):
            dx = np.abs(a[c].astype(np.float64) - want[c])
            tol = abs(kf) * np.abs(V[:, v] - W[:, v]) + 2 * np.spacing(np.abs(want[c]).astype(np.float32))
            assert np.all(dx <= tol + 1e-30), c
        before["s"] = {k: v.copy() for k, v in a.items()}

    P.subCycle_steps(TS, after_step=check)
    P.close()


def _read_cxx(path, n):
    raw = np.fromfile(path, dtype=np.uint8)
    half = raw.shape[0] // 2
    out = []
    for part in (raw[:half], raw[half:]):
        f = part[: 32 * n].view(np.float32).reshape(8, n)
        ids = part[32 * n: 40 * n].view(np.int64)
        mask = part[40 * n: 42 * n].view(np.uint16)
        d = dict(zip(("x", "y", "z", "vx", "vy", "vz", "phi", "mass"), f))
        d["id"], d["mask"] = ids, mask
        out.append(d)
    return out


def test_cxx_subcycle_dropin(tmp_path):
    """The C++ drop-in (ParticleSubCycle, reference caller pattern
    subCycle(gts)) equals its own map1/map2 steps and the Python front end on
    the same particles, bit for bit."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "hacc-coral_amd", "bin", "test_subcycle")
    n, g, nd = 30000, 20, 64
    res = []
    for mode in (0, 1):
        say(f"case C++ subcycle mode {mode}")
        out = tmp_path / f"sub{mode}.bin"
        subprocess.run([exe, str(n), str(g), str(nd), "77", str(mode), str(out)], check=True,
                       timeout=120)
        res.append(_read_cxx(out, n))
    (init, fin0), (_, fin1) = res
    for k in ARR:
        assert bits_equal(fin0[k], fin1[k]), k
    a = {k: v.copy() for k, v in init.items()}
    P = make((g, g, g), nd)
    P.set_particles(a["x"], a["y"], a["z"], a["vx"], a["vy"], a["vz"], a["mass"], a["phi"],
                    a["id"], a["mask"])
    P.subCycle(TS)
    P.close()
    for k in ARR:
        assert bits_equal(a[k], fin0[k]), k


def test_subcycle_from_restart_file(tmp_path):
    """Particles read from a restart file (serial layout) cycle exactly like
    the arrays they were written from (readRestart sets unit masses)."""
    from haccsr import restart as R
    ng, nd = (16, 16, 16), 64
    st = particles(20000, ng, 21)
    st["mass"][:] = 1.0
    say("case subcycle from a restart file")
    p = R.RestartParticles(**{k: st[k] for k in ("x", "y", "z", "vx", "vy", "vz", "phi")},
                           id=st["id"], mask=st["mask"], mass=st["mass"])
    f = tmp_path / "restart.bin"
    R.write_restart(f, p)
    q = R.read_restart(f)
    a = {k: v.copy() for k, v in st.items()}
    outs = []
    for src in (a, {k: getattr(q, k) for k in ARR}):
        P = make(ng, nd)
        P.set_particles(src["x"], src["y"], src["z"], src["vx"], src["vy"], src["vz"],
                        src["mass"], src["phi"], src["id"], src["mask"])
        P.subCycle(TS)
        P.close()
        outs.append(src)
    for k in ARR:
        assert bits_equal(outs[0][k], outs[1][k]), k
