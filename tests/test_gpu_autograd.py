"""Differentiable kick-drift on the GPU: haccsr.autograd.ShortRangeForce
(forward leaf kernel, backward hand-written adjoint with tree and lists
rebuilt from each sub-step's position checkpoint) over 5 sub-steps of
v += c F(x), x += dt v with dt = 0.01 and map2's kick coefficient
c = dt / (4 pi) (Particles.cxx:1231-1234 with gpscal = fscal = 1).

Checked against the FP64 oracle with the topology of each sub-step frozen
to the one the GPU used: the FP64 trajectory, the FP64 reverse-mode
gradient, and central finite differences (h = 1e-3, FP64) on 256
coordinates (32 on the larger case).  The gradient is compared on its force-dependent part
(dJ/dx0 - wx and dJ/dv0 - (wv + N dt wx)), which is what the adjoint
contributes.
"""
import numpy as np
import pytest
import torch

from haccsr import ics
from haccsr.autograd import ForceSolver, Images, kick_drift
import kdref as K
from oracle import oracle as O

pytestmark = pytest.mark.gpu

NSUB, DT = 5, 0.01
KICK = DT / (4 * np.pi)
OVERLOAD = 3.2


def say(msg):
    print(msg, flush=True)


def rel_rms(a, b):
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))


def make_case(n_side, seed, nd, oa):
    box = float(n_side)
    x, y, z = ics.clustered(n_side, box, max(1, round(500 * (n_side / 256) ** 3)), seed)
    load = ics.add_ghosts(x, y, z, box, OVERLOAD)
    base = np.stack([x, y, z], 1).astype(np.float32)
    idx = load.alive_index
    full = np.stack([load.x, load.y, load.z], 1)
    shift = (full.astype(np.float64) - base[idx].astype(np.float64)).astype(np.float32)
    rng = np.random.default_rng(seed)
    v0 = rng.normal(size=base.shape).astype(np.float32)
    wx = rng.normal(size=base.shape)
    wv = rng.normal(size=base.shape)
    # tree box with a margin for the drift; force box = the box's own region
    lo = np.full(3, -1.0, np.float32)
    hi = np.full(3, box + 2 * OVERLOAD + 1.0, np.float32)
    flo = np.full(3, OVERLOAD, np.float32)
    fhi = np.full(3, box + OVERLOAD, np.float32)
    tp = O.TreeParams(lo, hi, flo, fhi, nd=nd, oa=oa)
    return dict(tp=tp, base=base, idx=idx, shift=shift, v0=v0, wx=wx, wv=wv)


def run_gpu(c, scheme="kd"):
    tp = c["tp"]
    solver = ForceSolver(tp.min_loc, tp.max_loc, tp.min_force, tp.max_force,
                         nd=tp.nd, oa=tp.oa, fsrrmax=tp.fsrrmax, rsm=tp.rsm, record=True)
    n = c["base"].shape[0]
    img = Images(torch.from_numpy(c["idx"]).cuda(), torch.from_numpy(c["shift"]).cuda(), n)
    x0 = torch.from_numpy(c["base"]).cuda().requires_grad_(True)
    v0 = torch.from_numpy(c["v0"]).cuda().requires_grad_(True)
    xN, vN = kick_drift(x0, v0, solver, NSUB, DT, KICK, img, scheme=scheme)
    wx = torch.from_numpy(c["wx"].astype(np.float32)).cuda()
    wv = torch.from_numpy(c["wv"].astype(np.float32)).cuda()
    J = (xN * wx).sum() + (vN * wv).sum()
    J.backward()
    torch.cuda.synchronize()
    out = dict(xN=xN.detach().cpu().numpy().astype(np.float64),
               vN=vN.detach().cpu().numpy().astype(np.float64),
               gx=x0.grad.cpu().numpy().astype(np.float64),
               gv=v0.grad.cpu().numpy().astype(np.float64),
               topos=[K.topo_from_record(t) for t in solver.topologies])
    solver.ctx.close()
    return out


# (12, 32, 0.5): pseudo-particle heavy, 256 FD coordinates; (16, 128, 0.1): the
# leaf size and opening angle of the 128^3 differentiation load, 32 FD
# coordinates (its denser halo makes the h = 1e-3 truncation error larger).
@pytest.mark.parametrize("n_side,nd,oa,seed,n_fd,fd_tol", [(12, 32, 0.5, 41, 256, 1e-3),
                                                           (16, 128, 0.1, 42, 32, 5e-3)])
def test_kick_drift_gradient(n_side, nd, oa, seed, n_fd, fd_tol):
    say(f"case kick-drift n_side={n_side} nd={nd} oa={oa} seed={seed}")
    c = make_case(n_side, seed, nd, oa)
    g = run_gpu(c)
    assert len(g["topos"]) == NSUB
    tp, base, idx, shift = c["tp"], c["base"], c["idx"], c["shift"]
    args = (tp, g["topos"], base, c["v0"], idx, shift.astype(np.float64), DT, KICK)

    # forward: GPU trajectory vs FP64 trajectory on the same topologies
    _, x64, v64 = K.trajectory(*args)
    free_x = base + NSUB * DT * c["v0"].astype(np.float64)
    e_x = rel_rms(g["xN"] - free_x, x64 - free_x)
    e_v = rel_rms(g["vN"] - c["v0"], v64 - c["v0"])
    say(f"  forward rel rms (force part) x {e_x:.3e} v {e_v:.3e}")
    assert e_x < 1e-4 and e_v < 1e-4

    # gradient: GPU autograd vs FP64 reverse pass
    gx64, gv64 = K.gradient(*args, c["wx"], c["wv"])
    fx64, fv64 = gx64 - c["wx"], gv64 - (c["wv"] + NSUB * DT * c["wx"])
    fx, fv = g["gx"] - c["wx"], g["gv"] - (c["wv"] + NSUB * DT * c["wx"])
    e_gx, e_gv = rel_rms(fx, fx64), rel_rms(fv, fv64)
    say(f"  gradient rel rms (force part) dx0 {e_gx:.3e} dv0 {e_gv:.3e}  "
        f"|force part|/|wx| {np.sqrt(np.mean(fx64 ** 2)):.3e}")
    assert e_gx < 1e-3 and e_gv < 1e-3

    # finite differences (FP64, h = 1e-3) on n_fd coordinates of x0
    rng = np.random.default_rng(seed + 1)
    n = base.shape[0]
    coords = list(zip(rng.integers(0, n, n_fd).tolist(), rng.integers(0, 3, n_fd).tolist()))
    fd = K.fd_gradient(*args, c["wx"], c["wv"], coords, h=1e-3)
    w = np.array([c["wx"][i, a] for i, a in coords])
    fd_f = fd - w
    ref_f = np.array([fx64[i, a] for i, a in coords])
    gpu_f = np.array([fx[i, a] for i, a in coords])
    e_fd, e_fd_gpu = rel_rms(ref_f, fd_f), rel_rms(gpu_f, fd_f)
    say(f"  FD: FP64 reverse vs FD {e_fd:.3e}; GPU vs FD {e_fd_gpu:.3e}")
    assert e_fd < fd_tol and e_fd_gpu < 2 * fd_tol


def test_kick_drift_gradient_is_deterministic():
    say("case kick-drift determinism n_side=12")
    c = make_case(12, 43, 32, 0.5)
    a, b = run_gpu(c), run_gpu(c)
    assert np.array_equal(a["gx"], b["gx"]) and np.array_equal(a["gv"], b["gv"])
    assert np.array_equal(a["xN"], b["xN"])


def test_single_force_vjp_with_resident_tree():
    """One ShortRangeForce call: backward reuses the resident tree (no
    rebuild) and must equal the FP64 adjoint at the same positions."""
    say("case single vjp resident tree")
    c = make_case(12, 44, 32, 0.5)
    tp = c["tp"]
    solver = ForceSolver(tp.min_loc, tp.max_loc, tp.min_force, tp.max_force,
                         nd=tp.nd, oa=tp.oa, record=True)
    from haccsr.autograd import ShortRangeForce
    pos_np = (c["base"][c["idx"]] + c["shift"]).astype(np.float32)
    pos = torch.from_numpy(pos_np).cuda().requires_grad_(True)
    F = ShortRangeForce.apply(pos, solver)
    lam = np.random.default_rng(5).normal(size=pos_np.shape)
    (F * torch.from_numpy(lam.astype(np.float32)).cuda()).sum().backward()
    topo = K.topo_from_record(solver.topologies[0])
    lam_l = lam.astype(np.float32).astype(np.float64)
    ref = K.vjp_all(tp, topo, pos_np.astype(np.float64), lam_l)
    got = pos.grad.cpu().numpy().astype(np.float64)
    e = rel_rms(got, ref)
    say(f"  rel rms vs FP64 adjoint {e:.3e}")
    assert e < 1e-4
    solver.ctx.close()


def test_drift_kick_drift_gradient():
    """The drift-kick-drift form of Particles::subCycle through the same
    autograd path: GPU gradient vs the FP64 reverse pass and 64 FD
    coordinates."""
    say("case drift-kick-drift n_side=12")
    c = make_case(12, 45, 32, 0.5)
    g = run_gpu(c, scheme="dkd")
    tp, base, idx, shift = c["tp"], c["base"], c["idx"], c["shift"]
    args = (tp, g["topos"], base, c["v0"], idx, shift.astype(np.float64), DT, KICK)
    gx64, gv64 = K.gradient(*args, c["wx"], c["wv"], scheme="dkd")
    fx64, fv64 = gx64 - c["wx"], gv64 - (c["wv"] + NSUB * DT * c["wx"])
    fx, fv = g["gx"] - c["wx"], g["gv"] - (c["wv"] + NSUB * DT * c["wx"])
    e_gx, e_gv = rel_rms(fx, fx64), rel_rms(fv, fv64)
    rng = np.random.default_rng(46)
    coords = list(zip(rng.integers(0, base.shape[0], 64).tolist(), rng.integers(0, 3, 64).tolist()))
    fd = K.fd_gradient(*args, c["wx"], c["wv"], coords, h=1e-3, scheme="dkd")
    w = np.array([c["wx"][i, a] for i, a in coords])
    e_fd = rel_rms(np.array([fx[i, a] for i, a in coords]), fd - w)
    say(f"  gradient rel rms dx0 {e_gx:.3e} dv0 {e_gv:.3e}; GPU vs FD {e_fd:.3e}")
    assert e_gx < 1e-3 and e_gv < 1e-3 and e_fd < 2e-3
