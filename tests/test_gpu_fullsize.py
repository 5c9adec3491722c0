"""Parity at production shape: the 256^3 clustered load (16.8M particles +
1.2M ghosts, leaf size 512, tree depth 21, lists chunked into work items)
and the ensemble of 64 clustered boxes of 32^3 to 128^3 in one context.

Tree order, node table and interaction lists are compared bit for bit with
the oracle over the whole load; forces over random samples of listed leaves
(the oracle's float kernel over all of them would take minutes); the float
forces' error against the device FP64 evaluation of the same lists (equal
to the CPU FP64 oracle, tests/test_gpu_fp64.py) is stated and bounded."""
import numpy as np
import pytest

from haccsr import Context, ics
from oracle import oracle as O
from test_gpu_parity import check_lists, check_tree, listed_mask, rel_rms, say

pytestmark = pytest.mark.gpu

ND_B = 512          # the leaf size the autotuner picks for this load
SAMPLE_B = 2000     # listed leaves whose forces are checked against the oracle
ND_E = 256
SAMPLE_E = 10       # per box


def tree_params(load, nd):
    lo, hi = load.tree_box
    flo, fhi = load.force_box
    return O.TreeParams(np.asarray(lo, np.float32), np.asarray(hi, np.float32),
                        np.asarray(flo, np.float32), np.asarray(fhi, np.float32), nd=nd)


def sampled_forces(tp, tree, lists, leaf_idx):
    """The oracle's Step16 kernel over the lists of the chosen leaves (each
    leaf's list does not depend on the others)."""
    sub = O.build_lists(tp, tree, leaves=lists.leaves[leaf_idx])
    assert np.array_equal(sub.entries, lists.entries[leaf_idx])
    rx, ry, rz = O.forces_step16(tp, tree, sub)
    return sub, np.stack([rx, ry, rz], 1)


@pytest.mark.timeout(900)
def test_config_b_full_size():
    say("case config B 256^3 clustered, nd=512: load")
    load = ics.config_load("B")
    tp = tree_params(load, ND_B)
    say(f"  {load.n} particles; oracle tree and lists")
    tree = O.build_tree(tp, load.x, load.y, load.z)
    lists = O.build_lists(tp, tree)
    say("  device build and lists")
    ctx = Context(nd=ND_B)
    X, Y, Z = load.x.copy(), load.y.copy(), load.z.copy()
    M = np.ones(load.n, np.float32)
    ids = np.arange(load.n, dtype=np.int64)
    ctx.build(X, Y, Z, M, ids=ids, min_loc=load.tree_box[0], max_loc=load.tree_box[1])
    ctx.lists(*load.force_box)
    st = ctx.stats()
    say(f"  nodes {st.nodes} depth {st.max_depth} listed leaves {st.listed_leaves} "
        f"pairs {st.pairs}")
    assert st.max_depth >= 20 and st.pairs > 5e11
    check_tree(tree, ctx, (X, Y, Z, M, ids))
    check_lists(lists, ctx)
    say("  forces: device float and FP64, oracle on sampled leaves")
    ax, ay, az = ctx.forces()
    G = np.stack([ax, ay, az], 1)
    dx, dy, dz = ctx.forces_f64()
    sel = listed_mask(tree, lists)
    e64 = rel_rms(G[sel].astype(np.float64), np.stack([dx, dy, dz], 1)[sel])
    rng = np.random.default_rng(256)
    pick = np.sort(rng.choice(lists.leaves.shape[0], SAMPLE_B, replace=False))
    sub, R = sampled_forces(tp, tree, lists, pick)
    ssel = listed_mask(tree, sub)
    e = rel_rms(G[ssel], R[ssel])
    D = np.stack([dx, dy, dz], 1)
    eref = rel_rms(R[ssel].astype(np.float64), D[ssel])  # the reference kernel's own rounding
    say(f"  float vs oracle on {SAMPLE_B} leaves ({int(ssel.sum())} particles): rel-RMS {e:.3e} "
        f"(oracle float kernel vs FP64 there: {eref:.3e}); device float vs FP64 over all "
        f"{int(sel.sum())} listed particles: rel-RMS {e64:.3e}")
    assert e <= 1e-5
    assert e64 <= 1e-5
    ctx.close()


@pytest.mark.timeout(900)
def test_config_e_64_boxes():
    say("case config E: 64 clustered boxes 32^3..128^3, nd=256: loads")
    loads = ics.ensemble_loads()
    assert len(loads) == 64
    offs = np.cumsum([0] + [b.n for b in loads]).astype(np.int64)
    X = np.concatenate([b.x for b in loads])
    Y = np.concatenate([b.y for b in loads])
    Z = np.concatenate([b.z for b in loads])
    M = np.ones(X.shape[0], np.float32)
    ids = np.arange(X.shape[0], dtype=np.int64)
    lo = np.stack([b.tree_box[0] for b in loads]).astype(np.float32)
    hi = np.stack([b.tree_box[1] for b in loads]).astype(np.float32)
    flo = np.stack([b.force_box[0] for b in loads]).astype(np.float32)
    fhi = np.stack([b.force_box[1] for b in loads]).astype(np.float32)
    say(f"  {X.shape[0]} particles; device build and lists of all boxes")
    ctx = Context(nd=ND_E)
    ctx.build_many(offs, X, Y, Z, M, lo, hi, ids=ids)
    ctx.lists_many(flo, fhi)
    ax, ay, az = ctx.forces()
    dx, dy, dz = ctx.forces_f64()
    nodes = ctx.nodes()
    leaves, seg_ptr, sa, sb, ent = ctx.get_lists()
    listed = np.zeros(X.shape[0], bool)
    for l in leaves:
        listed[nodes["offset"][l]:nodes["offset"][l] + nodes["count"][l]] = True
    e64 = rel_rms(np.stack([ax, ay, az], 1)[listed].astype(np.float64),
                  np.stack([dx, dy, dz], 1)[listed])
    say(f"  device float vs FP64 over all {int(listed.sum())} listed particles: rel-RMS {e64:.3e}")
    assert e64 <= 1e-5
    rng = np.random.default_rng(64)
    base = lk = sk = 0
    worst = 0.0
    for b, load in enumerate(loads):
        tp = tree_params(load, ND_E)
        tree = O.build_tree(tp, load.x, load.y, load.z)
        lists = O.build_lists(tp, tree)
        a, e = offs[b], offs[b + 1]
        assert np.array_equal(ids[a:e] - a, tree.id), f"box {b} permutation"
        for arr, ref in ((X, tree.x), (Y, tree.y), (Z, tree.z)):
            assert np.array_equal(arr[a:e].view(np.uint32), ref.view(np.uint32)), f"box {b} positions"
        nn = tree.nodes.shape[0]
        mine = nodes[base:base + nn]
        for f in ("count", "cl", "cr", "ppm", "tdr", "xmin", "xmax", "xc"):
            got, ref = mine[f].copy(), tree.nodes[f]
            if f in ("cl", "cr"):
                got = np.where(got > 0, got - base, 0)
            if got.dtype.kind == "f":
                got, ref = got.view(np.uint32), ref.view(np.uint32)
            assert np.array_equal(got, ref), f"box {b} node field {f}"
        assert np.array_equal(mine["offset"] - np.where(mine["count"] > 0, a, 0),
                              tree.nodes["offset"]), f"box {b} offsets"
        nl = lists.leaves.shape[0]
        assert np.array_equal(leaves[lk:lk + nl] - base, lists.leaves), f"box {b} leaves"
        assert np.array_equal(ent[lk:lk + nl], lists.entries), f"box {b} list lengths"
        ns = lists.seg_a.shape[0]
        s0 = seg_ptr[lk]
        assert np.array_equal(seg_ptr[lk:lk + nl + 1] - s0, lists.seg_ptr), f"box {b} seg_ptr"
        # segments are particle ranges (offset, count; offsets shifted by the
        # box's first particle) or pseudo-particles (node id, -1; ids shifted
        # by the box's first node)
        assert sk == s0
        ga, gb = sa[s0:s0 + ns], sb[s0:s0 + ns]
        assert np.array_equal(gb, lists.seg_b), f"box {b} segment counts"
        shift = np.where(gb >= 0, a, base)
        assert np.array_equal(ga - shift, lists.seg_a), f"box {b} segment starts"
        k = min(SAMPLE_E, nl)
        pick = np.sort(rng.choice(nl, k, replace=False))
        sub, R = sampled_forces(tp, tree, lists, pick)
        ssel = listed_mask(tree, sub)
        Gb = np.stack([ax[a:e], ay[a:e], az[a:e]], 1)
        err = rel_rms(Gb[ssel], R[ssel])
        worst = max(worst, err)
        assert err <= 1e-5, f"box {b} forces {err:.3e}"
        base += nn
        lk += nl
        sk += ns
        if b % 16 == 15:
            say(f"  {b + 1}/64 boxes bit-exact, worst sampled force rel-RMS {worst:.3e}")
    assert lk == leaves.shape[0]
    ctx.close()
