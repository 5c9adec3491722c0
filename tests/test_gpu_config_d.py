"""Differentiation at production shape: the 128^3 clustered load (seed
1004, 62 NFW halos, 3.2 ghost shell: 2.36M particles), five short-range
kick-drift sub-steps v += c F(x), x += dt v (dt = 0.01, c = dt / (4 pi)) through
haccsr.autograd (forward leaf kernel, backward hand-written adjoint with the
tree and lists rebuilt from each sub-step's position checkpoint), with a
random Gaussian cotangent on the final positions and velocities.

The FP64 reference is the device FP64 oracle (hsr_forces_f64_at,
hsr_adjoint_f64_at; equal to the CPU FP64 oracle to 1e-15,
tests/test_gpu_fp64.py) on the topology each sub-step used, rebuilt from its
recorded float positions:
  * the FP64 trajectory, against the GPU's;
  * the FP64 reverse-mode gradient linearised at the positions the GPU used,
    against the GPU's (its force-dependent part, dJ/dx0 - wx and
    dJ/dv0 - (wv + N dt wx)); and the one along the FP64 trajectory, which
    adds the float trajectory's rounding;
  * central finite differences of the FP64 objective (h = 1e-3) on 16 single
    coordinates of x0, and along 4 random unit directions over 256
    coordinates each (one objective difference checks a combination of 256
    gradient components).
The CPU FP64 oracle would take about 90 s per force and 1000 s per adjoint
at this size, hence the device oracle."""
import numpy as np
import pytest
import torch

from haccsr import Context, ics
from haccsr.autograd import ForceSolver, Images, kick_drift

pytestmark = pytest.mark.gpu

NSUB, DT = 5, 0.01
KICK = DT / (4 * np.pi)
OVERLOAD = 3.2
N_SIDE, ND, OA, SEED = 128, 128, 0.1, 1004
N_FD, N_DIR, DIR_COORDS, H = 16, 4, 256, 1e-3


def say(msg):
    print(msg, flush=True)


def rel_rms(a, b):
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))


class Oracle64:
    """FP64 force and VJP of one sub-step's frozen topology on the device."""

    def __init__(self, topo, lo, hi, flo, fhi):
        p = topo.pos
        n = p.shape[0]
        self.ctx = Context(nd=ND, oa=OA)
        x, y, z = (np.ascontiguousarray(p[:, k]) for k in range(3))
        ids = np.arange(n, dtype=np.int64)
        self.ctx.build(x, y, z, np.ones(n, np.float32), ids=ids, min_loc=lo, max_loc=hi)
        self.ctx.lists(flo, fhi)
        assert np.array_equal(ids, topo.perm), "rebuilt topology differs"
        nodes = self.ctx.nodes()
        for f in ("count", "offset", "cl", "cr"):
            assert np.array_equal(nodes[f], topo.nodes[f]), f
        self.perm = topo.perm

    def _dev(self, a):
        return [torch.from_numpy(np.ascontiguousarray(a[self.perm, k])).cuda() for k in range(3)]

    def _host(self, ts):
        out = np.empty((self.perm.shape[0], 3))
        out[self.perm] = np.stack([t.cpu().numpy() for t in ts], 1)
        return out

    def force(self, pos):
        return self._host(self.ctx.forces_f64_at(*self._dev(pos)))

    def vjp(self, pos, lam):
        return self._host(self.ctx.adjoint_f64_at(*self._dev(pos), *self._dev(lam)))


def expand(x, idx, shift):
    return x[idx] + shift


def trajectory(orcs, x0, v0, idx, shift):
    x, v = x0.astype(np.float64).copy(), v0.astype(np.float64).copy()
    n = x.shape[0]
    xs = []
    for o in orcs:
        xs.append(x.copy())
        v = v + KICK * o.force(expand(x, idx, shift))[:n]
        x = x + DT * v
    return xs, x, v


def objective(orcs, x0, v0, idx, shift, wx, wv):
    _, x, v = trajectory(orcs, x0, v0, idx, shift)
    return float(np.sum(wx * x) + np.sum(wv * v))


def gradient(orcs, positions, n, idx, wx, wv):
    """Reverse pass for J = wx . x_N + wv . v_N, the force of sub-step k
    linearised at positions[k] (every particle, images included)."""
    gx, gv = wx.copy(), wv.copy()
    for o, pos in zip(reversed(orcs), reversed(positions)):
        gv = gv + DT * gx
        lam = np.zeros((idx.shape[0], 3))
        lam[:n] = KICK * gv
        g_all = o.vjp(pos, lam)
        gx = gx + np.stack([np.bincount(idx, weights=g_all[:, k], minlength=n)
                            for k in range(3)], 1)
    return gx, gv


@pytest.mark.timeout(900)
def test_config_d_kick_drift_gradient():
    say(f"case config D {N_SIDE}^3 clustered, nd={ND}, oa={OA}: load")
    box = float(N_SIDE)
    x, y, z = ics.clustered(N_SIDE, box, max(1, round(500 * (N_SIDE / 256) ** 3)), SEED)
    load = ics.add_ghosts(x, y, z, box, OVERLOAD)
    base = np.stack([x, y, z], 1).astype(np.float32)
    idx = load.alive_index
    full = np.stack([load.x, load.y, load.z], 1)
    shift = (full.astype(np.float64) - base[idx].astype(np.float64)).astype(np.float32)
    shift64 = shift.astype(np.float64)  # the images the GPU applies
    rng = np.random.default_rng(SEED)
    v0 = rng.normal(size=base.shape).astype(np.float32)
    wx = rng.normal(size=base.shape)
    wv = rng.normal(size=base.shape)
    lo = np.full(3, -1.0, np.float32)
    hi = np.full(3, box + 2 * OVERLOAD + 1.0, np.float32)
    flo = np.full(3, OVERLOAD, np.float32)
    fhi = np.full(3, box + OVERLOAD, np.float32)
    say(f"  {load.n} particles ({base.shape[0]} alive); GPU autograd over {NSUB} sub-steps")
    solver = ForceSolver(lo, hi, flo, fhi, nd=ND, oa=OA, record=True)
    img = Images(torch.from_numpy(idx).cuda(), torch.from_numpy(shift).cuda(), base.shape[0])
    X0 = torch.from_numpy(base).cuda().requires_grad_(True)
    V0 = torch.from_numpy(v0).cuda().requires_grad_(True)
    xN, vN = kick_drift(X0, V0, solver, NSUB, DT, KICK, img)
    J = (xN * torch.from_numpy(wx.astype(np.float32)).cuda()).sum() + \
        (vN * torch.from_numpy(wv.astype(np.float32)).cuda()).sum()
    J.backward()
    torch.cuda.synchronize()
    gx32 = X0.grad.cpu().numpy().astype(np.float64)
    gv32 = V0.grad.cpu().numpy().astype(np.float64)
    xN32 = xN.detach().cpu().numpy().astype(np.float64)
    vN32 = vN.detach().cpu().numpy().astype(np.float64)
    assert len(solver.topologies) == NSUB
    pairs = solver.ctx.stats().pairs
    solver.ctx.close()
    say(f"  device FP64 oracle on the {NSUB} recorded topologies ({pairs:.3e} pairs at the last)")
    orcs = [Oracle64(t, lo, hi, flo, fhi) for t in solver.topologies]

    # forward: the GPU trajectory against the FP64 one (force-dependent part)
    _, x64, v64 = trajectory(orcs, base, v0, idx, shift64)
    free_x = base + NSUB * DT * v0.astype(np.float64)
    e_x = rel_rms(xN32 - free_x, x64 - free_x)
    e_v = rel_rms(vN32 - v0, v64 - v0)
    say(f"  forward rel rms (force part): x {e_x:.3e}, v {e_v:.3e}")

    # gradient: GPU autograd against the FP64 reverse pass linearised at the
    # positions the GPU used (the float trajectory), which isolates the
    # adjoint's arithmetic; and against the FP64 reverse pass along the FP64
    # trajectory, which adds the float trajectory's own rounding (positions
    # near 131 carry 7.6e-6 of rounding per sub-step, and the force varies on
    # the 0.1 softening length inside the halos)
    n = base.shape[0]
    xs64, _, _ = trajectory(orcs, base, v0, idx, shift64)
    gx_t, gv_t = gradient(orcs, [t.pos.astype(np.float64) for t in solver.topologies], n, idx, wx, wv)
    gx64, gv64 = gradient(orcs, [expand(xk, idx, shift64) for xk in xs64], n, idx, wx, wv)
    fx_t, fv_t = gx_t - wx, gv_t - (wv + NSUB * DT * wx)
    fx64, fv64 = gx64 - wx, gv64 - (wv + NSUB * DT * wx)
    fx32, fv32 = gx32 - wx, gv32 - (wv + NSUB * DT * wx)
    e_gx, e_gv = rel_rms(fx32, fx_t), rel_rms(fv32, fv_t)
    e_tx, e_tv = rel_rms(fx32, fx64), rel_rms(fv32, fv64)
    say(f"  gradient rel rms (force part) vs FP64 on the GPU trajectory: dJ/dx0 {e_gx:.3e}, "
        f"dJ/dv0 {e_gv:.3e}; vs FP64 on the FP64 trajectory: {e_tx:.3e}, {e_tv:.3e}; "
        f"|force part| / |wx| = {np.sqrt(np.mean(fx64 ** 2)):.3e}")

    # finite differences of the FP64 objective, single coordinates
    coords = list(zip(rng.integers(0, n, N_FD).tolist(), rng.integers(0, 3, N_FD).tolist()))
    fd = np.zeros(N_FD)
    for k, (i, a) in enumerate(coords):
        xp, xm = base.astype(np.float64).copy(), base.astype(np.float64).copy()
        xp[i, a] += H
        xm[i, a] -= H
        fd[k] = (objective(orcs, xp, v0, idx, shift64, wx, wv) -
                 objective(orcs, xm, v0, idx, shift64, wx, wv)) / (2 * H)
    w = np.array([wx[i, a] for i, a in coords])
    ref = np.array([fx64[i, a] for i, a in coords])
    got = np.array([fx32[i, a] for i, a in coords])
    e_fd, e_fd32 = rel_rms(ref, fd - w), rel_rms(got, fd - w)
    say(f"  FD on {N_FD} coordinates (h = {H}): FP64 reverse vs FD {e_fd:.3e}, GPU vs FD {e_fd32:.3e}")

    # finite differences along random unit directions over DIR_COORDS
    # coordinates each (components +-1/16, step H: one objective difference
    # checks a combination of DIR_COORDS gradient components)
    rows = []
    for k in range(N_DIR):
        d = np.zeros_like(base, np.float64)
        pick = rng.choice(n * 3, DIR_COORDS, replace=False)
        d.reshape(-1)[pick] = rng.choice([-1.0, 1.0], DIR_COORDS) / np.sqrt(DIR_COORDS)
        fdd = (objective(orcs, base + H * d, v0, idx, shift64, wx, wv) -
               objective(orcs, base - H * d, v0, idx, shift64, wx, wv)) / (2 * H)
        rows.append((fdd - np.sum(wx * d), np.sum(fx64 * d), np.sum(fx32 * d)))
    rows = np.array(rows)
    e_dir = rel_rms(rows[:, 1], rows[:, 0])
    e_dir32 = rel_rms(rows[:, 2], rows[:, 0])
    say(f"  FD along {N_DIR} unit directions over {DIR_COORDS} coordinates: FP64 reverse vs FD "
        f"{e_dir:.3e}, GPU vs FD {e_dir32:.3e}")
    for o in orcs:
        o.ctx.close()
    assert e_x < 1e-3 and e_v < 1e-3
    assert e_gx < 1e-3 and e_gv < 1e-3
    assert e_tx < 2e-2 and e_tv < 2e-2
    assert e_fd < 5e-3 and e_fd32 < 1e-2
    assert e_dir < 5e-3 and e_dir32 < 1e-2
