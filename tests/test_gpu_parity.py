"""GPU parity: tree order, node table and interaction lists bit-exact with
the oracle; forces within 1e-5 relative RMS of the reference's float kernel
(the oracle's restatement of nbody1 with the Step16 law)."""
import sys

import numpy as np
import pytest

from haccsr import Context, _lib, ics
from oracle import oracle as O

pytestmark = pytest.mark.gpu

NODE_FIELDS = ["count", "offset", "cl", "cr", "ppm", "tdr", "xmin", "xmax", "xc"]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run_both(x, y, z, m, lo, hi, flo, fhi, nd, fsrrmax=3.12, rsm=0.1, oa=0.1):
    tp = O.TreeParams(np.asarray(lo, np.float32), np.asarray(hi, np.float32),
                      np.asarray(flo, np.float32), np.asarray(fhi, np.float32),
                      fsrrmax=fsrrmax, rsm=rsm, oa=oa, nd=nd)
    tree = O.build_tree(tp, x, y, z, m)
    lists = O.build_lists(tp, tree)
    ctx = Context(nd=nd, fsrrmax=fsrrmax, rsm=rsm, oa=oa)
    X, Y, Z, M = x.copy(), y.copy(), z.copy(), m.copy()
    ids = np.arange(x.shape[0], dtype=np.int64)
    ctx.build(X, Y, Z, M, ids=ids, min_loc=lo, max_loc=hi)
    ctx.lists(flo, fhi)
    return tp, tree, lists, ctx, (X, Y, Z, M, ids)


def check_tree(tree, ctx, arrays):
    X, Y, Z, M, ids = arrays
    assert np.array_equal(ids, tree.id), "permutation differs"
    assert np.array_equal(X.view(np.uint32), tree.x.view(np.uint32))
    assert np.array_equal(Y.view(np.uint32), tree.y.view(np.uint32))
    assert np.array_equal(Z.view(np.uint32), tree.z.view(np.uint32))
    nodes = ctx.nodes()
    assert nodes.shape == tree.nodes.shape, (nodes.shape, tree.nodes.shape)
    for f in NODE_FIELDS:
        a, b = nodes[f], tree.nodes[f]
        if a.dtype.kind == "f":
            a, b = a.view(np.uint32), b.view(np.uint32)
        bad = np.nonzero(np.any((a != b).reshape(a.shape[0], -1), axis=1))[0]
        assert bad.size == 0, f"node field {f} differs at {bad[:5]}"


def check_lists(lists, ctx):
    leaves, seg_ptr, sa, sb, ent = ctx.get_lists()
    assert np.array_equal(leaves, lists.leaves)
    assert np.array_equal(seg_ptr, lists.seg_ptr)
    assert np.array_equal(sa, lists.seg_a)
    assert np.array_equal(sb, lists.seg_b)
    assert np.array_equal(ent, lists.entries)


def rel_rms(a, b):
    d = np.sum((a - b) ** 2)
    s = np.sum(b ** 2)
    return float(np.sqrt(d / s)) if s > 0 else float(np.sqrt(d))


def listed_mask(tree, lists):
    sel = np.zeros(tree.x.shape[0], bool)
    for l in lists.leaves:
        nd = tree.nodes[l]
        sel[nd["offset"]:nd["offset"] + nd["count"]] = True
    return sel


def check_forces(tp, tree, lists, ctx, tol=1e-5):
    """hsr_forces returns the Step16 acceleration a_i = sum_j m_j f d
    (BGQStep16.c:167-187); the reference's generic nbody1 accumulates
    m_i * a_i per pair (RCBForceTree.cxx:610).  Both oracle restatements are
    checked: the Step16 one directly, nbody1 after dividing by m_i."""
    ax, ay, az = ctx.forces()
    rx, ry, rz = O.forces_step16(tp, tree, lists)
    sel = listed_mask(tree, lists)
    G = np.stack([ax, ay, az], 1)[sel]
    R = np.stack([rx, ry, rz], 1)[sel]
    nx, ny, nz = O.forces_f32(tp, tree, lists)
    Nb = (np.stack([nx, ny, nz], 1) / tree.mass[:, None])[sel]
    assert rel_rms(G, Nb) <= tol, f"vs nbody1/m_i: {rel_rms(G, Nb):.3e}"
    e = rel_rms(G, R)
    assert np.all(np.stack([ax, ay, az], 1)[~sel] == 0)
    assert e <= tol, f"relative RMS {e:.3e} > {tol}"
    return e


def uniform(n, box, seed):
    rng = np.random.default_rng(seed)
    p = (rng.random((n, 3)) * box).astype(np.float32)
    return p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), np.ones(n, np.float32)


@pytest.mark.parametrize("n,box,nd,seed", [(3000, 8.0, 16, 1), (20000, 16.0, 64, 2)])
def test_uniform_small(n, box, nd, seed):
    say(f"case uniform n={n} box={box} nd={nd} seed={seed}")
    x, y, z, m = uniform(n, box, seed)
    lo, hi = [0, 0, 0], [box] * 3
    flo, fhi = [1.0] * 3, [box - 1.0] * 3
    tp, tree, lists, ctx, arrs = run_both(x, y, z, m, lo, hi, flo, fhi, nd)
    check_tree(tree, ctx, arrs)
    check_lists(lists, ctx)
    e = check_forces(tp, tree, lists, ctx)
    say(f"  rel rms {e:.3e}")


def test_config_a_parity():
    say("case config A 64^3 plummer mix, nd=128")
    load = ics.config_load("A")
    lo, hi = load.tree_box
    flo, fhi = load.force_box
    m = np.ones(load.n, np.float32)
    tp, tree, lists, ctx, arrs = run_both(load.x, load.y, load.z, m, lo, hi, flo, fhi, 128)
    check_tree(tree, ctx, arrs)
    check_lists(lists, ctx)
    e = check_forces(tp, tree, lists, ctx)
    st = ctx.stats()
    say(f"  nodes {st.nodes} leaves {st.listed_leaves} pairs {st.pairs} rel rms {e:.3e} "
        f"build {st.build_ms:.2f} ms lists {st.lists_ms:.2f} ms force {st.force_ms:.2f} ms")


def test_clustered_nd192_multipass():
    say("case clustered 48^3 (config-B generator), nd=192")
    load = ics.config_load("B", n_side=48)
    lo, hi = load.tree_box
    flo, fhi = load.force_box
    m = np.ones(load.n, np.float32)
    tp, tree, lists, ctx, arrs = run_both(load.x, load.y, load.z, m, lo, hi, flo, fhi, 192)
    check_tree(tree, ctx, arrs)
    check_lists(lists, ctx)
    e = check_forces(tp, tree, lists, ctx)
    say(f"  rel rms {e:.3e} depth {ctx.stats().max_depth}")


def test_chunked_lists_and_determinism():
    say("case chunked lists (chunk=256) and bitwise determinism")
    x, y, z, m = uniform(30000, 14.0, 3)
    lo, hi = [0, 0, 0], [14.0] * 3
    flo, fhi = [1.0] * 3, [13.0] * 3
    tp, tree, lists, ctx, arrs = run_both(x, y, z, m, lo, hi, flo, fhi, 96)
    e1 = check_forces(tp, tree, lists, ctx)
    a1 = np.stack(ctx.forces(), 1)
    a2 = np.stack(ctx.forces(), 1)
    assert np.array_equal(a1.view(np.uint32), a2.view(np.uint32)), "not bitwise reproducible"
    ctx.set_tuning(chunk=256)
    ctx.lists(flo, fhi)
    e2 = check_forces(tp, tree, lists, ctx)
    b1 = np.stack(ctx.forces(), 1)
    assert np.array_equal(b1.view(np.uint32), np.stack(ctx.forces(), 1).view(np.uint32))
    say(f"  rel rms unchunked {e1:.3e} chunked {e2:.3e}")


@pytest.mark.parametrize("law,n_interp", [(O.LAW_SR_POLY5, 0), (O.LAW_SR_POLY6, 0),
                                          (O.LAW_SR_FIT, 0), (O.LAW_SR_INTERP, 1024),
                                          (O.LAW_SR_INTERP, 97), (O.LAW_SR_INTERP, 2)])
def test_sr_laws_match_oracle(law, n_interp):
    """ForceLawSR over each grid-force evaluator (ForceLaw.cxx:39-192): the
    5th- and 6th-order polynomials, FGridEvalFit (FGrid::fgor per pair; the
    reference's default law) and FGridEvalInterp with 1024, 97 and 2 table
    points (the table from the C++ drop-in's FGrid::fgor_r2_interp): the
    device kernel against the oracle's nbody1 with the same law, per unit
    target mass, on unequal masses; the tree and lists bit-exact as with the
    default law."""
    from haccsr import FGrid, FGridEvalInterp
    say(f"case SR law {law} n_interp={n_interp}")
    rng = np.random.default_rng(40 + law)
    n, box, nd = 20000, 14.0, 64
    p = (rng.random((n, 3)) * box).astype(np.float32)
    x, y, z = p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy()
    m = rng.uniform(0.5, 2.0, n).astype(np.float32)
    lo, hi = np.zeros(3, np.float32), np.full(3, box, np.float32)
    flo, fhi = np.full(3, 1.0, np.float32), np.full(3, box - 1.0, np.float32)
    rmax = 3.116326355  # FGrid::rmax (ForceLaw.cxx:32), the drop-in's fsrrmax
    tp = O.TreeParams(lo, hi, flo, fhi, nd=nd, fsrrmax=rmax, n_interp=n_interp)
    tree = O.build_tree(tp, x, y, z, m)
    lists = O.build_lists(tp, tree)
    table = FGridEvalInterp(FGrid(), n_interp).table() if n_interp else None
    ctx = Context(nd=nd, law=law, fsrrmax=rmax, interp_table=table)
    ids = np.arange(n, dtype=np.int64)
    X, Y, Z, M = x.copy(), y.copy(), z.copy(), m.copy()
    ctx.build(X, Y, Z, M, ids=ids, min_loc=lo, max_loc=hi)
    ctx.lists(flo, fhi)
    check_tree(tree, ctx, (X, Y, Z, M, ids))
    check_lists(lists, ctx)
    ax, ay, az = ctx.forces()
    sel = listed_mask(tree, lists)
    rx, ry, rz = O.forces_f32(tp, tree, lists, law=law)
    ref = (np.stack([rx, ry, rz], 1) / tree.mass[:, None])[sel]
    e = rel_rms(np.stack([ax, ay, az], 1)[sel], ref)
    say(f"  rel rms {e:.3e}")
    assert e <= 1e-5
    ctx.close()


def test_newton_law_refused_on_device():
    """ForceLawNewton is not implemented on the device: refused at context
    creation (no CPU fallback).  An interpolation context without its table
    refuses to compute forces, and the adjoint and FP64 paths refuse the
    laws they do not implement."""
    with pytest.raises(_lib.HsrError):
        Context(nd=64, law=_lib.LAW_NEWTON)
    with pytest.raises(_lib.HsrError, match="interp_table"):
        Context(nd=64, law=_lib.LAW_SR_INTERP)
    x, y, z, m = uniform(3000, 8.0, 9)
    ctx = Context(nd=32, law=_lib.LAW_SR_FIT)
    ctx.build(x, y, z, m, min_loc=(0, 0, 0), max_loc=(8, 8, 8))
    ctx.lists((1, 1, 1), (7, 7, 7))
    ctx.forces()
    with pytest.raises(_lib.HsrError, match="polynomial laws"):
        ctx.adjoint(*[np.zeros(3000, np.float32)] * 3)
    with pytest.raises(_lib.HsrError, match="polynomial laws"):
        ctx.forces_f64()
    ctx.close()


def test_tuning_validation_and_chunk_replan():
    """hsr_set_tuning rejects out-of-range fields and keeps the old tuning;
    a new chunk size re-plans lists already built, so the next force phase
    (no new hsr_lists) equals the one after a fresh list build bitwise."""
    import ctypes
    say("case tuning validation and re-plan on a chunk change")
    x, y, z, m = uniform(30000, 14.0, 5)
    flo, fhi = [1.0] * 3, [13.0] * 3
    tp, tree, lists, ctx, arrs = run_both(x, y, z, m, [0, 0, 0], [14.0] * 3, flo, fhi, 96)
    before = _lib.Tuning()
    assert ctx.L.hsr_get_tuning(ctx.h, ctypes.byref(before)) == 0
    for bad in (dict(tile=64), dict(i_per_lane=3), dict(unroll=3), dict(tile=128, i_per_lane=4),
                dict(chunk=-1), dict(chunk=(1 << 30) + 1),
                dict(waves_per_cu=-2), dict(waves_per_cu=1000), dict(force_pass=96)):
        with pytest.raises(_lib.HsrError):
            ctx.set_tuning(**bad)
        now = _lib.Tuning()
        assert ctx.L.hsr_get_tuning(ctx.h, ctypes.byref(now)) == 0
        assert bytes(now) == bytes(before), bad
    ctx.set_tuning(tile=256, i_per_lane=2, waves_per_cu=40)
    e1 = check_forces(tp, tree, lists, ctx)
    ctx.set_tuning(chunk=256)
    replanned = np.stack(ctx.forces(), 1).view(np.uint32).copy()
    ctx.lists(flo, fhi)
    fresh = np.stack(ctx.forces(), 1).view(np.uint32)
    assert np.array_equal(replanned, fresh)
    e2 = check_forces(tp, tree, lists, ctx)
    say(f"  rel rms chunk 8192 {e1:.3e}, chunk 256 {e2:.3e}")


SHAPES = [(2, 2, 256), (2, 1, 256), (2, 4, 256), (1, 2, 256), (1, 4, 256), (4, 1, 256),
          (2, 2, 128), (2, 2, 512)]


def test_every_built_kernel_shape_matches_oracle():
    """Each leaf-kernel shape the tuner can pick (i-particles per lane, j
    entries per unrolled step, entries per LDS tile) gives forces within the
    tolerance of the oracle on a clustered load with long, chunked lists;
    shapes change only the order of the float sums."""
    say("case every built kernel shape, clustered 40^3, nd=256")
    load = ics.config_load("B", n_side=40)
    m = np.ones(load.n, np.float32)
    tp, tree, lists, ctx, arrs = run_both(load.x, load.y, load.z, m, *load.tree_box,
                                          *load.force_box, 256)
    check_lists(lists, ctx)
    ctx.set_tuning(chunk=1024)
    base = None
    for r, u, t in SHAPES:
        ctx.set_tuning(i_per_lane=r, unroll=u, tile=t, chunk=1024)
        e = check_forces(tp, tree, lists, ctx)
        st = ctx.stats()
        f = np.stack(ctx.forces(), 1)
        if base is None:
            base = f
        d = rel_rms(f, base)
        say(f"  shape {(r, u, t)}: rel rms vs oracle {e:.3e}, vs default shape {d:.3e}, "
            f"unmasked {st.unmasked_pairs / max(st.evaluated_pairs, 1):.3f} of evaluated")
        assert d <= 2e-6
    ctx.close()


def test_duplicates_and_degenerate_nodes():
    say("case duplicate positions (unsplittable nodes)")
    x, y, z, m = uniform(4000, 10.0, 4)
    dup = np.full(300, 5.25, np.float32)
    x = np.concatenate([x, dup]); y = np.concatenate([y, dup + 0.5]); z = np.concatenate([z, dup - 0.5])
    m = np.ones(x.shape[0], np.float32)
    tp, tree, lists, ctx, arrs = run_both(x, y, z, m, [0, 0, 0], [10.0] * 3, [1.0] * 3, [9.0] * 3, 32)
    check_tree(tree, ctx, arrs)
    check_lists(lists, ctx)
    check_forces(tp, tree, lists, ctx)
    st = ctx.stats()
    assert st.empty_leaves > 0 and st.max_ppn >= 300


@pytest.mark.parametrize("n", [0, 1, 2, 17])
def test_tiny_inputs(n):
    say(f"case tiny n={n}")
    x, y, z, m = uniform(n, 4.0, 5)
    tp, tree, lists, ctx, arrs = run_both(x, y, z, m, [0, 0, 0], [4.0] * 3, [0.5] * 3, [3.5] * 3, 8)
    if n:
        check_tree(tree, ctx, arrs)
        check_lists(lists, ctx)
        ax, ay, az = ctx.forces()
        rx, ry, rz = O.forces_f32(tp, tree, lists)
        assert np.allclose(ax, rx, rtol=1e-5, atol=1e-6)


def test_mixed_magnitudes_and_masses():
    say("case mixed coordinate magnitudes and non-unit masses (exact cm paths)")
    rng = np.random.default_rng(12)
    n = 20000
    p = (rng.random((n, 3)) * 30.0).astype(np.float32)
    p[: n // 10] *= np.float32(1e-4)          # coordinates near 0 force rounding
    m = (0.5 + rng.random(n)).astype(np.float32)
    tp, tree, lists, ctx, arrs = run_both(p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), m,
                                          [0, 0, 0], [30.0] * 3, [1.0] * 3, [29.0] * 3, 48)
    check_tree(tree, ctx, arrs)
    check_lists(lists, ctx)
    check_forces(tp, tree, lists, ctx)


@pytest.mark.parametrize("case", ["negative", "big_nodes"])
def test_exact_sum_paths(case):
    """The centre-of-mass sums on every path: negative coordinates (no
    integer sum: the ordered paths), big nodes (> 2048 particles, the block
    tables) whose sums are integers in units of the finest lsb, and big
    nodes whose sums round (coordinates near 0 and masses with full
    mantissas)."""
    rng = np.random.default_rng(31 if case == "negative" else 32)
    n = 60000
    if case == "negative":
        say("case exact-sum paths: coordinates in [-16, 16), unit masses")
        p = (rng.random((n, 3)) * 32.0 - 16.0).astype(np.float32)
        m = np.ones(n, np.float32)
        lo, hi, flo, fhi = [-16.0] * 3, [16.0] * 3, [-15.0] * 3, [15.0] * 3
    else:
        say("case exact-sum paths: big nodes, integer sums and rounding sums")
        p = (rng.random((n, 3)) * 40.0).astype(np.float32)
        p[: n // 3] = np.round(p[: n // 3] * 64.0) / np.float32(64.0)  # coarse lsb
        p[n // 3: n // 3 + 200] *= np.float32(1e-6)                     # fine lsb near 0
        m = np.where(rng.random(n) < 0.5, np.float32(1.0),
                     (0.5 + rng.random(n)).astype(np.float32)).astype(np.float32)
        lo, hi, flo, fhi = [0.0] * 3, [40.0] * 3, [1.0] * 3, [39.0] * 3
    tp, tree, lists, ctx, arrs = run_both(p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), m,
                                          lo, hi, flo, fhi, 64)
    check_tree(tree, ctx, arrs)
    check_lists(lists, ctx)
    check_forces(tp, tree, lists, ctx)


def test_ensemble_boxes_one_launch():
    """Independent boxes in one context (hsr_build_many / hsr_lists_many):
    every box's tree, lists and forces equal the oracle run on that box
    alone; one persistent launch covers all boxes' lists."""
    say("case ensemble of 6 boxes")
    rng = np.random.default_rng(21)
    boxes = []
    for b in range(6):
        ns = int(rng.integers(12, 22))
        load = ics.add_ghosts(*ics.plummer_mix(ns ** 3, float(ns), 1100 + b), float(ns), 3.2)
        load, _ = ics.cell_sort(load)
        boxes.append(load)
    offs = np.cumsum([0] + [b.n for b in boxes]).astype(np.int64)
    X = np.concatenate([b.x for b in boxes]); Y = np.concatenate([b.y for b in boxes])
    Z = np.concatenate([b.z for b in boxes]); M = np.ones(X.shape[0], np.float32)
    ids = np.arange(X.shape[0], dtype=np.int64)
    lo = np.stack([b.tree_box[0] for b in boxes]); hi = np.stack([b.tree_box[1] for b in boxes])
    flo = np.stack([b.force_box[0] for b in boxes]); fhi = np.stack([b.force_box[1] for b in boxes])
    ctx = Context(nd=64)
    ctx.build_many(offs, X, Y, Z, M, lo, hi, ids=ids)
    ctx.lists_many(flo, fhi)
    ax, ay, az = ctx.forces()
    nodes = ctx.nodes()
    leaves, seg_ptr, sa, sb, ent = ctx.get_lists()
    base = 0
    lk = 0
    for b, load in enumerate(boxes):
        tp = O.TreeParams(lo[b], hi[b], flo[b], fhi[b], nd=64)
        tree = O.build_tree(tp, load.x, load.y, load.z)
        lists = O.build_lists(tp, tree)
        a, e = offs[b], offs[b + 1]
        assert np.array_equal(ids[a:e] - a, tree.id), f"box {b} permutation"
        nn = tree.nodes.shape[0]
        mine = nodes[base:base + nn]
        for f in ("count", "cl", "cr", "ppm", "tdr", "xmin", "xmax", "xc"):
            got, ref = mine[f].copy(), tree.nodes[f]
            if f in ("cl", "cr"):
                got = np.where(got > 0, got - base, 0)
            if got.dtype.kind == "f":
                got, ref = got.view(np.uint32), ref.view(np.uint32)
            assert np.array_equal(got, ref), f"box {b} node field {f}"
        assert np.array_equal(mine["offset"] - np.where(mine["count"] > 0, a, 0), tree.nodes["offset"])
        nl = lists.leaves.shape[0]
        assert np.array_equal(leaves[lk:lk + nl] - base, lists.leaves), f"box {b} leaves"
        assert np.array_equal(ent[lk:lk + nl], lists.entries), f"box {b} list lengths"
        rx, ry, rz = O.forces_step16(tp, tree, lists)
        sel = listed_mask(tree, lists)
        G = np.stack([ax[a:e], ay[a:e], az[a:e]], 1)[sel]
        R = np.stack([rx, ry, rz], 1)[sel]
        assert rel_rms(G, R) <= 1e-5, f"box {b} forces {rel_rms(G, R):.3e}"
        base += nn
        lk += nl
    say(f"  {len(boxes)} boxes, {ctx.stats().listed_leaves} listed leaves, pairs {ctx.stats().pairs}")


def test_graph_replay_matches_eager():
    """The force phase captured as a hipGraph: every replay, with eager work
    and synchronisations between replays, equals the eager launch bitwise."""
    say("case hipGraph replay of the force phase")
    x, y, z, m = uniform(40000, 16.0, 9)
    tp, tree, lists, ctx, arrs = run_both(x, y, z, m, [0, 0, 0], [16.0] * 3, [1.0] * 3, [15.0] * 3, 96)
    eager = np.stack(ctx.forces(), 1).view(np.uint32).copy()
    ctx.set_tuning(use_graph=1)
    for rep in range(4):
        got = np.stack(ctx.forces(), 1).view(np.uint32)
        assert np.array_equal(got, eager), f"replay {rep} differs"
        ctx.set_tuning(use_graph=0)
        assert np.array_equal(np.stack(ctx.forces(), 1).view(np.uint32), eager)
        ctx.set_tuning(use_graph=1)
    # a new list build invalidates the captured graph
    ctx.set_tuning(chunk=512)
    ctx.lists([1.0] * 3, [15.0] * 3)
    e2 = check_forces(tp, tree, lists, ctx)
    say(f"  replays ok; after re-listing rel rms {e2:.3e}")


@pytest.mark.parametrize("hops", [1, 2, 7])
def test_partition_chains_by_pointer_jumping(hops):
    """The partition's swap chains that take more than chain_hops stretch
    steps are resolved by the pointer-jumping pass: with 1, 2 and 7 steps
    (the default is 256, which almost every chain needs fewer of) most
    chains go through it, and the tree, permutation and lists stay bit-exact
    with the oracle."""
    say(f"case partition chains by pointer jumping, chain_hops {hops}")
    for (x, y, z, m), box, nd in ((uniform(30000, 14.0, 21), 14.0, 64),
                                  (uniform(5000, 6.0, 22), 6.0, 16)):
        lo, hi, flo, fhi = [0, 0, 0], [box] * 3, [1.0] * 3, [box - 1.0] * 3
        tp = O.TreeParams(np.asarray(lo, np.float32), np.asarray(hi, np.float32),
                          np.asarray(flo, np.float32), np.asarray(fhi, np.float32),
                          fsrrmax=3.12, rsm=0.1, oa=0.1, nd=nd)
        tree = O.build_tree(tp, x, y, z, m)
        lists = O.build_lists(tp, tree)
        ctx = Context(nd=nd, fsrrmax=3.12, rsm=0.1, oa=0.1)
        ctx.set_tuning(chain_hops=hops)
        X, Y, Z, M = x.copy(), y.copy(), z.copy(), m.copy()
        ids = np.arange(x.shape[0], dtype=np.int64)
        ctx.build(X, Y, Z, M, ids=ids, min_loc=lo, max_loc=hi)
        ctx.lists(flo, fhi)
        check_tree(tree, ctx, (X, Y, Z, M, ids))
        check_lists(lists, ctx)
        ctx.close()


def test_chain_hops_validation():
    ctx = Context(nd=64, fsrrmax=3.12, rsm=0.1, oa=0.1)
    for bad in (-1, 4097):
        with pytest.raises(_lib.HsrError, match="chain_hops"):
            ctx.set_tuning(chain_hops=bad)
    ctx.set_tuning(chain_hops=4096)
    ctx.close()


def test_graph_replayed_across_list_builds():
    """The captured force phase is replayed, not captured again, after a new
    build and list pass of the same particle set (fixed grids, device-side
    counts, buffer addresses unchanged), and every replay equals the eager
    force phase bitwise; moved particles (new tree, new lists) replay too
    when their buffers fit, and still equal the eager result and the
    oracle."""
    say("case force-phase graph across list builds")
    x, y, z, m = uniform(40000, 16.0, 11)
    lo, hi, flo, fhi = [0, 0, 0], [16.0] * 3, [1.0] * 3, [15.0] * 3
    tp, tree, lists, ctx, arrs = run_both(x, y, z, m, lo, hi, flo, fhi, 96)
    eager = np.stack(ctx.forces(), 1).view(np.uint32).copy()
    ctx.set_tuning(use_graph=1)
    first = np.stack(ctx.forces(), 1).view(np.uint32).copy()
    assert np.array_equal(first, eager)
    s0 = ctx.stats()
    for rep in range(3):
        X, Y, Z, M = x.copy(), y.copy(), z.copy(), m.copy()
        ctx.build(X, Y, Z, M, min_loc=lo, max_loc=hi)
        ctx.lists(flo, fhi)
        got = np.stack(ctx.forces(), 1).view(np.uint32)
        assert np.array_equal(got, eager), f"replay {rep} after a rebuild differs"
    s1 = ctx.stats()
    say(f"  captures {s0.graph_captures} -> {s1.graph_captures}, replays {s0.graph_replays} -> {s1.graph_replays}")
    assert s1.graph_captures == s0.graph_captures, "a rebuild of the same particles captured again"
    assert s1.graph_replays == s0.graph_replays + 3
    # moved particles: new tree and lists, results as eager and as the oracle
    rng = np.random.default_rng(5)
    x2 = np.clip(x + rng.normal(0, 0.05, x.shape).astype(np.float32), 0, np.float32(15.999))
    X2, Y2, Z2, M2 = x2.copy(), y.copy(), z.copy(), m.copy()
    ctx.build(X2, Y2, Z2, M2, min_loc=lo, max_loc=hi)
    ctx.lists(flo, fhi)
    g2 = np.stack(ctx.forces(), 1).view(np.uint32).copy()
    ctx.set_tuning(use_graph=0)
    e2 = np.stack(ctx.forces(), 1).view(np.uint32)
    assert np.array_equal(g2, e2)
    tree2 = O.build_tree(tp, x2.copy(), y.copy(), z.copy(), m.copy())
    lists2 = O.build_lists(tp, tree2)
    err = check_forces(tp, tree2, lists2, ctx)
    say(f"  moved particles: graph = eager bitwise, rel rms vs oracle {err:.3e}")


def test_culling_and_target_order_change_rounding_only():
    """Skipping list entries beyond the cutoff of a pass's targets, and
    sorting a leaf's targets along its longest edge, leave every pair that
    the cutoff keeps: forces agree with the unculled tree-order kernel to
    float rounding of the reordered sums (5e-6 relative RMS, below the
    1e-5 bar each setting meets against the oracle), and the evaluated pair
    counts order as expected."""
    say("case culling / target order on clustered 48^3, nd=256")
    load = ics.config_load("B", n_side=48)
    lo, hi = load.tree_box
    flo, fhi = load.force_box
    m = np.ones(load.n, np.float32)
    tp, tree, lists, ctx, arrs = run_both(load.x, load.y, load.z, m, lo, hi, flo, fhi, 256)
    out, ev = {}, {}
    for cull_off, tree_order in ((1, 1), (0, 1), (0, 0)):
        ctx.set_tuning(cull_off=cull_off, tree_order_targets=tree_order)
        out[(cull_off, tree_order)] = np.stack(ctx.forces(), 1)
        ev[(cull_off, tree_order)] = ctx.stats().evaluated_pairs
        check_forces(tp, tree, lists, ctx)
    ref = out[(1, 1)]
    for k in ((0, 1), (0, 0)):
        assert rel_rms(out[k], ref) < 5e-6, (k, rel_rms(out[k], ref))
    pairs = ctx.stats().pairs
    say(f"  evaluated pairs: all {ev[(1, 1)]}, leaf-box culled {ev[(0, 1)]}, "
        f"sorted targets {ev[(0, 0)]} of {pairs}")
    assert ev[(1, 1)] == pairs and ev[(0, 0)] <= ev[(0, 1)] < ev[(1, 1)]
