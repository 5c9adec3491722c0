"""GPU parity of the quadrupole tree (TDPTS = 12, RCBQuadrupoleForceTree):
node table, the twelve pseudo masses per node, tree order and lists
bit-exact with the oracle; forces within 1e-5 relative RMS of the oracle's
float kernel; the device FP64 evaluation equal to the CPU FP64 oracle."""
import numpy as np
import pytest

from haccsr import Context, HsrError, ics
from oracle import oracle as O
from test_gpu_parity import check_forces, check_lists, check_tree, rel_rms, say, uniform

pytestmark = pytest.mark.gpu


def run_q(x, y, z, m, lo, hi, flo, fhi, nd, oa=0.1):
    tp = O.TreeParams(np.asarray(lo, np.float32), np.asarray(hi, np.float32),
                      np.asarray(flo, np.float32), np.asarray(fhi, np.float32),
                      oa=oa, nd=nd, tdpts=12)
    tree = O.build_tree(tp, x, y, z, m)
    lists = O.build_lists(tp, tree)
    ctx = Context(nd=nd, oa=oa, tdpts=12)
    X, Y, Z, M = x.copy(), y.copy(), z.copy(), m.copy()
    ids = np.arange(x.shape[0], dtype=np.int64)
    ctx.build(X, Y, Z, M, ids=ids, min_loc=lo, max_loc=hi)
    ctx.lists(flo, fhi)
    return tp, tree, lists, ctx, (X, Y, Z, M, ids)


def check_ppm(tree, ctx):
    got = ctx.pseudo_masses()
    assert got.shape == tree.ppm_t.shape
    bad = np.nonzero(np.any(got.view(np.uint32) != tree.ppm_t.view(np.uint32), axis=1))[0]
    assert bad.size == 0, f"pseudo masses differ at nodes {bad[:5]}"
    assert np.any(got != 0)


@pytest.mark.parametrize("n,box,nd,oa,seed", [(3000, 8.0, 16, 0.5, 1), (20000, 16.0, 64, 0.8, 2)])
def test_quadrupole_uniform(n, box, nd, oa, seed):
    say(f"case quadrupole uniform n={n} nd={nd} oa={oa}")
    x, y, z, _ = uniform(n, box, seed)
    m = np.random.default_rng(seed).uniform(0.5, 2.0, n).astype(np.float32)
    lo, hi = [0, 0, 0], [box] * 3
    flo, fhi = [1.0] * 3, [box - 1.0] * 3
    tp, tree, lists, ctx, arrs = run_q(x, y, z, m, lo, hi, flo, fhi, nd, oa)
    check_tree(tree, ctx, arrs)
    check_ppm(tree, ctx)
    check_lists(lists, ctx)
    assert np.any(lists.seg_b < 0)
    e = check_forces(tp, tree, lists, ctx)
    say(f"  rel rms {e:.3e}, pseudo segments {int(np.sum(lists.seg_b < 0))}")


def test_quadrupole_clustered_and_fp64():
    say("case quadrupole clustered 48^3, nd=128")
    load = ics.config_load("B", n_side=48)
    lo, hi = load.tree_box
    flo, fhi = load.force_box
    m = np.ones(load.n, np.float32)
    tp, tree, lists, ctx, arrs = run_q(load.x, load.y, load.z, m, lo, hi, flo, fhi, 128, 0.5)
    check_tree(tree, ctx, arrs)
    check_ppm(tree, ctx)
    check_lists(lists, ctx)
    e = check_forces(tp, tree, lists, ctx)
    rx, ry, rz = O.forces_f64(tp, tree, lists)
    gx, gy, gz = ctx.forces_f64()
    ef = rel_rms(np.stack([gx, gy, gz], 1), np.stack([rx, ry, rz], 1))
    say(f"  rel rms {e:.3e}; device FP64 vs CPU FP64 {ef:.3e}")
    assert ef < 1e-12


def test_quadrupole_degenerate_nodes():
    say("case quadrupole duplicates (unsplittable nodes)")
    rng = np.random.default_rng(5)
    n = 6000
    p = (rng.random((n, 3)) * 10.0).astype(np.float32)
    p[:700] = p[0]                 # one unsplittable clump of 700
    p[700:760] = p[1]              # and a small one
    m = np.ones(n, np.float32)
    tp, tree, lists, ctx, arrs = run_q(p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), m,
                                       [0, 0, 0], [10.0] * 3, [1.0] * 3, [9.0] * 3, 32, 0.5)
    check_tree(tree, ctx, arrs)
    check_ppm(tree, ctx)
    check_lists(lists, ctx)
    check_forces(tp, tree, lists, ctx)


def test_quadrupole_ensemble_matches_single_boxes():
    say("case quadrupole ensemble of 3 boxes")
    boxes, outs = [], []
    rng = np.random.default_rng(8)
    for b, (n, L) in enumerate(((4000, 10.0), (9000, 14.0), (2500, 8.0))):
        x, y, z, _ = uniform(n, L, 40 + b)
        boxes.append((x, y, z, L))
    X = np.concatenate([b[0] for b in boxes]); Y = np.concatenate([b[1] for b in boxes])
    Z = np.concatenate([b[2] for b in boxes]); M = np.ones(X.shape[0], np.float32)
    off = np.cumsum([0] + [b[0].shape[0] for b in boxes])
    lo = np.zeros((3, 3), np.float32); hi = np.array([[b[3]] * 3 for b in boxes], np.float32)
    flo = np.ones((3, 3), np.float32); fhi = hi - 1
    ctx = Context(nd=32, oa=0.5, tdpts=12)
    ctx.build_many(off, X, Y, Z, M, lo, hi)
    ctx.lists_many(flo, fhi)
    fx, fy, fz = ctx.forces()
    for b, (x, y, z, L) in enumerate(boxes):
        c1 = Context(nd=32, oa=0.5, tdpts=12)
        xs, ys, zs, ms = x.copy(), y.copy(), z.copy(), np.ones(x.shape[0], np.float32)
        c1.build(xs, ys, zs, ms, min_loc=[0, 0, 0], max_loc=[L] * 3)
        c1.lists([1.0] * 3, [L - 1] * 3)
        gx, gy, gz = c1.forces()
        sl = slice(off[b], off[b + 1])
        assert np.array_equal(X[sl].view(np.uint32), xs.view(np.uint32))
        assert rel_rms(np.stack([fx[sl], fy[sl], fz[sl]], 1), np.stack([gx, gy, gz], 1)) < 1e-6
        c1.close()
    ctx.close()


def test_quadrupole_adjoint_is_refused():
    x, y, z, m = uniform(2000, 8.0, 3)
    ctx = Context(nd=16, tdpts=12)
    ctx.build(x, y, z, m, min_loc=[0, 0, 0], max_loc=[8.0] * 3)
    ctx.lists([1.0] * 3, [7.0] * 3)
    lam = np.ones(2000, np.float32)
    with pytest.raises(HsrError, match="monopole"):
        ctx.adjoint(lam, lam, lam)
    ctx.close()
