"""GPU adjoint (hsr_adjoint) against the FP64 oracle adjoint on the same
tree and lists; the oracle adjoint itself is pinned by finite differences
(tests/test_oracle.py)."""
import numpy as np
import pytest

from haccsr import Context, ics
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def setup(x, y, z, box, nd, oa, margin=1.0, seed=3):
    m = np.ones_like(x)
    lo, hi = np.zeros(3, np.float32), np.full(3, box, np.float32)
    flo, fhi = np.full(3, margin, np.float32), np.full(3, box - margin, np.float32)
    tp = O.TreeParams(lo, hi, flo, fhi, nd=nd, oa=oa)
    tree = O.build_tree(tp, x, y, z, m)
    lists = O.build_lists(tp, tree)
    ctx = Context(nd=nd, oa=oa)
    X, Y, Z, M = x.copy(), y.copy(), z.copy(), m.copy()
    ctx.build(X, Y, Z, M, min_loc=lo, max_loc=hi)
    ctx.lists(flo, fhi)
    rng = np.random.default_rng(seed)
    lam = rng.normal(size=(x.shape[0], 3))
    sel = np.zeros(x.shape[0], bool)
    for l in lists.leaves:
        nd_ = tree.nodes[l]
        sel[nd_["offset"]:nd_["offset"] + nd_["count"]] = True
    lam[~sel] = 0.0
    return tp, tree, lists, ctx, lam


def rel_rms(a, b):
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))


# adjoint_mode 1 = fused schedule (one pair evaluation for both sides),
# 2 = separate target and source passes
@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("n,box,nd,oa,chunk", [(6000, 10.0, 16, 0.7, 0), (20000, 14.0, 64, 0.1, 0),
                                               (20000, 14.0, 64, 0.5, 300), (30000, 12.0, 256, 0.5, 0)])
def test_adjoint_matches_oracle(n, box, nd, oa, chunk, mode):
    say(f"case adjoint n={n} nd={nd} oa={oa} chunk={chunk} mode={mode}")
    rng = np.random.default_rng(n)
    p = (rng.random((n, 3)) * box).astype(np.float32)
    tp, tree, lists, ctx, lam = setup(p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), box, nd, oa)
    ctx.set_tuning(adjoint_mode=mode)
    if chunk:
        ctx.set_tuning(chunk=chunk)
        ctx.lists(tp.min_force, tp.max_force)
    ref = O.adjoint_f64(tp, tree, lists, lam)
    L32 = lam.astype(np.float32)
    gx, gy, gz = ctx.adjoint(L32[:, 0].copy(), L32[:, 1].copy(), L32[:, 2].copy())
    got = np.stack([gx, gy, gz], 1)
    e = rel_rms(got, ref)
    say(f"  rel rms vs FP64 {e:.3e}  pseudo segments {(lists.seg_b < 0).sum()}")
    assert e < 1e-4
    # bitwise reproducible
    g2 = np.stack(ctx.adjoint(L32[:, 0].copy(), L32[:, 1].copy(), L32[:, 2].copy()), 1)
    assert np.array_equal(got.view(np.uint32), g2.view(np.uint32))


@pytest.mark.parametrize("mode", [1, 2])
def test_adjoint_config_a(mode):
    say(f"case adjoint config A (64^3, nd=128) mode={mode}")
    load = ics.config_load("A")
    lo, hi = load.tree_box
    flo, fhi = load.force_box
    m = np.ones(load.n, np.float32)
    tp = O.TreeParams(lo, hi, flo, fhi, nd=128)
    tree = O.build_tree(tp, load.x, load.y, load.z, m)
    lists = O.build_lists(tp, tree)
    ctx = Context(nd=128)
    X, Y, Z, M = load.x.copy(), load.y.copy(), load.z.copy(), m.copy()
    ctx.build(X, Y, Z, M, min_loc=lo, max_loc=hi)
    ctx.lists(flo, fhi)
    ctx.set_tuning(adjoint_mode=mode)
    rng = np.random.default_rng(1)
    lam = rng.normal(size=(load.n, 3))
    sel = np.zeros(load.n, bool)
    for l in lists.leaves:
        nd_ = tree.nodes[l]
        sel[nd_["offset"]:nd_["offset"] + nd_["count"]] = True
    lam[~sel] = 0.0
    ref = O.adjoint_f64(tp, tree, lists, lam)
    L32 = lam.astype(np.float32)
    got = np.stack(ctx.adjoint(L32[:, 0].copy(), L32[:, 1].copy(), L32[:, 2].copy()), 1)
    e = rel_rms(got, ref)
    ctx.forces()
    fwd = ctx.stats().kernel_ms
    ctx.adjoint(L32[:, 0].copy(), L32[:, 1].copy(), L32[:, 2].copy())
    adj = ctx.stats().adjoint_ms
    say(f"  rel rms vs FP64 {e:.3e}; forward kernel {fwd:.2f} ms, adjoint {adj:.2f} ms "
        f"(ratio {adj / fwd:.2f})")
    assert e < 1e-4


def test_adjoint_schedules_agree_on_large_leaves():
    """Leaves above 128 particles take several passes over their targets in
    both schedules (nd 300); fused and two-pass agree with the FP64 oracle
    and with each other to rounding."""
    say("case adjoint schedules nd=300")
    rng = np.random.default_rng(77)
    n, box = 20000, 12.0
    p = (rng.random((n, 3)) * box).astype(np.float32)
    tp, tree, lists, ctx, lam = setup(p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), box, 300, 0.5)
    assert tree.nodes["count"][lists.leaves].max() > 256
    L32 = lam.astype(np.float32)
    args = (L32[:, 0].copy(), L32[:, 1].copy(), L32[:, 2].copy())
    ref = O.adjoint_f64(tp, tree, lists, lam)
    out = {}
    for mode in (1, 2):
        ctx.set_tuning(adjoint_mode=mode)
        out[mode] = np.stack(ctx.adjoint(*args), 1).astype(np.float64)
        e = rel_rms(out[mode], ref)
        say(f"  mode {mode}: rel rms vs FP64 {e:.3e}")
        assert e < 1e-4
    assert rel_rms(out[1], out[2]) < 1e-5


@pytest.mark.parametrize("mode", [1, 2])
def test_adjoint_degenerate_tree(mode):
    """Duplicate positions make an unsplittable leaf of 200 particles (two
    passes over its targets in the fused schedule) next to ordinary leaves."""
    say(f"case adjoint degenerate duplicates mode={mode}")
    rng = np.random.default_rng(4)
    p = (rng.random((4000, 3)) * 10.0).astype(np.float32)
    dup = np.full(200, 5.25, np.float32)
    x = np.concatenate([p[:, 0], dup])
    y = np.concatenate([p[:, 1], dup + 0.5])
    z = np.concatenate([p[:, 2], dup - 0.5])
    tp, tree, lists, ctx, lam = setup(x, y, z, 10.0, 32, 0.5)
    assert tree.nodes["count"][lists.leaves].max() == 200
    ctx.set_tuning(adjoint_mode=mode)
    ref = O.adjoint_f64(tp, tree, lists, lam)
    L32 = lam.astype(np.float32)
    got = np.stack(ctx.adjoint(L32[:, 0].copy(), L32[:, 1].copy(), L32[:, 2].copy()), 1)
    e = rel_rms(got, ref)
    say(f"  rel rms vs FP64 {e:.3e}")
    assert e < 1e-4


def test_adjoint_graph_replay_matches_eager():
    """With tuning.use_graph the adjoint passes are captured once per plan and
    replayed: every replay (with eager work in between) equals the eager
    result bitwise, and a new list build re-captures."""
    say("case adjoint graph replay")
    rng = np.random.default_rng(31)
    n, box = 20000, 14.0
    p = (rng.random((n, 3)) * box).astype(np.float32)
    tp, tree, lists, ctx, lam = setup(p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), box, 64, 0.5)
    L32 = lam.astype(np.float32)
    args = (L32[:, 0].copy(), L32[:, 1].copy(), L32[:, 2].copy())
    eager = np.stack(ctx.adjoint(*args), 1)
    ctx.set_tuning(use_graph=1)
    for rep in range(4):
        got = np.stack(ctx.adjoint(*args), 1)
        assert np.array_equal(got.view(np.uint32), eager.view(np.uint32)), f"replay {rep}"
        ctx.forces()                      # eager work between replays
    ctx.lists(tp.min_force, tp.max_force)  # new plan -> new capture
    again = np.stack(ctx.adjoint(*args), 1)
    assert np.array_equal(again.view(np.uint32), eager.view(np.uint32))
    ref = O.adjoint_f64(tp, tree, lists, lam)
    assert rel_rms(eager.astype(np.float64), ref) < 1e-4
