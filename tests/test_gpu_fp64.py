"""FP64 evaluations on the device (hsr_forces_f64, hsr_leaf_momentum): the
device oracle against the CPU FP64 oracle on the same lists, the float
kernel's error against it, and the momentum stability check."""
import numpy as np
import pytest

from haccsr import Context, ics
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def build_both(x, y, z, m, lo, hi, flo, fhi, nd, oa):
    tp = O.TreeParams(np.asarray(lo, np.float32), np.asarray(hi, np.float32),
                      np.asarray(flo, np.float32), np.asarray(fhi, np.float32), nd=nd, oa=oa)
    tree = O.build_tree(tp, x, y, z, m)
    lists = O.build_lists(tp, tree)
    ctx = Context(nd=nd, oa=oa)
    X, Y, Z, M = x.copy(), y.copy(), z.copy(), m.copy()
    ctx.build(X, Y, Z, M, min_loc=tp.min_loc, max_loc=tp.max_loc)
    ctx.lists(tp.min_force, tp.max_force)
    return tp, tree, lists, ctx


def test_device_fp64_matches_cpu_fp64_oracle():
    say("case fp64 config A")
    load = ics.config_load("A")
    lo, hi = load.tree_box
    flo, fhi = load.force_box
    m = np.random.default_rng(2).uniform(0.5, 2.0, load.n).astype(np.float32)
    tp, tree, lists, ctx = build_both(load.x, load.y, load.z, m, lo, hi, flo, fhi, 128, 0.1)
    rx, ry, rz, rmom = O.forces_f64(tp, tree, lists, with_momentum=True)
    ax, ay, az, mom = ctx.forces_f64(with_momentum=True)
    ref = np.stack([rx, ry, rz], 1)
    got = np.stack([ax, ay, az], 1)
    e = np.sqrt(np.sum((got - ref) ** 2) / np.sum(ref ** 2))
    em = np.abs(mom - rmom).max() / np.abs(rmom).max()
    say(f"  device vs CPU FP64: rel rms {e:.3e}, leaf momentum max rel {em:.3e}")
    assert e < 1e-12 and em < 1e-9
    # the float kernel against the FP64 oracle (the accuracy target is 1e-5)
    fx, fy, fz = ctx.forces()
    f32 = np.stack([fx, fy, fz], 1).astype(np.float64)
    e32 = np.sqrt(np.sum((f32 - got) ** 2) / np.sum(got ** 2))
    say(f"  float kernel vs FP64: rel rms {e32:.3e}, max abs {np.abs(f32 - got).max():.3e}")
    assert e32 < 1e-5
    # FP64 momentum sums of the float forces: same as summing the returned floats
    lm = ctx.leaf_momentum(fx, fy, fz)
    sel_nodes = tree.nodes[lists.leaves]
    want = np.zeros_like(lm)
    for k, nd_ in enumerate(sel_nodes):
        a, b = nd_["offset"], nd_["offset"] + nd_["count"]
        want[k] = (tree.mass[a:b, None].astype(np.float64) * f32[a:b]).sum(0)
    assert np.allclose(lm, want, rtol=1e-12, atol=1e-9 * np.abs(want).max())


def test_momentum_stability_direct_lists():
    """With oa = 0 every interaction is a particle pair and, with the force box
    covering the tree, every pair appears in both directions: the FP64 total
    momentum of the FP64 forces vanishes to rounding, and that of the float
    forces to float rounding."""
    say("case momentum oa=0")
    rng = np.random.default_rng(9)
    n, box = 30000, 16.0
    p = (rng.random((n, 3)) * box).astype(np.float32)
    m = rng.uniform(0.5, 2.0, n).astype(np.float32)
    tp, tree, lists, ctx = build_both(p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), m,
                                      [0, 0, 0], [box] * 3, [-1.0] * 3, [box + 1.0] * 3, 64, 0.0)
    assert np.all(lists.seg_b >= 0) and lists.leaves.shape[0] > 0
    ax, ay, az, mom = ctx.forces_f64(with_momentum=True)
    a64 = np.stack([ax, ay, az], 1)
    scale = np.sum(np.abs(tree.mass[:, None] * a64))
    tot64 = np.abs(mom.sum(0)).max() / scale
    fx, fy, fz = ctx.forces()
    tot32 = np.abs(ctx.leaf_momentum(fx, fy, fz).sum(0)).max() / scale
    say(f"  |sum m a| / sum |m a|: FP64 forces {tot64:.3e}, float forces {tot32:.3e}")
    assert tot64 < 1e-12 and tot32 < 1e-6


def test_fp64_at_positions_matches_cpu_oracle():
    """hsr_forces_f64_at / hsr_adjoint_f64_at (the FP64 oracle of the
    differentiation checks: frozen tree and lists, double positions,
    pseudo-particles at the FP64 centre of mass of their members) against the
    CPU oracle's hsro_force_f64d / hsro_adjoint_f64 on the same topology, at
    positions moved off the float build positions, with unequal masses and
    a wide opening angle (many pseudo-particles)."""
    import torch
    say("case fp64 at positions, clustered 14^3, oa 0.5")
    load = ics.config_load("B", n_side=14)
    lo, hi = load.tree_box
    flo, fhi = load.force_box
    rng = np.random.default_rng(12)
    m = rng.uniform(0.5, 2.0, load.n).astype(np.float32)
    tp, tree, lists, ctx = build_both(load.x, load.y, load.z, m, lo, hi, flo, fhi, 32, 0.5)
    assert np.sum(lists.seg_b < 0) > 100, "case must contain pseudo-particles"
    pos = [a.astype(np.float64) + rng.normal(0, 1e-3, a.shape[0]) for a in (tree.x, tree.y, tree.z)]
    lam = rng.normal(size=(load.n, 3))
    cuda = [torch.from_numpy(np.ascontiguousarray(v)).cuda() for v in pos]
    lc = [torch.from_numpy(np.ascontiguousarray(lam[:, k])).cuda() for k in range(3)]
    f = np.stack([t.cpu().numpy() for t in ctx.forces_f64_at(*cuda)], 1)
    g = np.stack([t.cpu().numpy() for t in ctx.adjoint_f64_at(*cuda, *lc)], 1)
    mass = tree.mass.astype(np.float64)
    fr = O.force_f64_at(tp, tree.nodes, lists, pos, mass)
    gr = O.adjoint_f64_at(tp, tree.nodes, lists, pos, mass, lam)
    ef = np.sqrt(np.sum((f - fr) ** 2) / np.sum(fr ** 2))
    eg = np.sqrt(np.sum((g - gr) ** 2) / np.sum(gr ** 2))
    say(f"  device vs CPU FP64 at positions: force rel rms {ef:.3e}, VJP rel rms {eg:.3e}")
    assert ef < 1e-12 and eg < 1e-11
    ctx.close()
