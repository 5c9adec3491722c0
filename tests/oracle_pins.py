"""Independent checks that pin the oracle's tree walk decisions.

Test infrastructure, used by tests/test_oracle_lists.py (the oracle must pass
every check) and tests/test_oracle_mutations.py (each mistake planted in a
copy of the oracle must fail one).  None of these checks calls the oracle's
own walk: they restate what the reference's walk must produce in another
form and compare.

  check_leaf_order     calcInternodeForces (RCBForceTree.cxx:1151-1197):
                       reachable leaves whose box has, in every dimension, a
                       face strictly inside the force box, in right-first DFS
                       order = decreasing particle offset.
  check_list_order     calcInternodeForce (RCBForceTree.cxx:924-1144): the
                       LIFO walk pushes left then right, so every list is its
                       node ranges in strictly decreasing offset order (right
                       siblings of the leaf's ancestors top-down, then left
                       siblings bottom-up), then the leaf itself.
  check_list_walk      the same lists from a recursive walk whose per-node
                       decisions (opening test with its four space diagonals,
                       far cut, L-infinity gap) are evaluated for all nodes
                       at once in float32 numpy.
  check_pptdr          tdr = ppc * (smallest distance from xc to a face of
                       the node's box) (RCBForceTree.cxx:519-523, 787-797).
  check_law_constants  the oracle's polynomial coefficients are the float
                       values of the decimal literals in BGQStep16.c:167 and
                       ForceLaw.cxx:96-116 (tests/golden/ref_poly_coeffs.json).
"""
from __future__ import annotations

import json
import os

import numpy as np

from oracle import oracle as O

F = np.float32
HERE = os.path.dirname(os.path.abspath(__file__))


def reachable(nodes) -> np.ndarray:
    """Nodes reachable from the root through cl/cr (orphan slots of
    degenerate nodes are not)."""
    seen = np.zeros(nodes.shape[0], bool)
    stack = [0]
    while stack:
        v = stack.pop()
        seen[v] = True
        for c in (nodes["cl"][v], nodes["cr"][v]):
            if c > 0:
                stack.append(int(c))
    return seen


def expected_leaf_order(tp, nodes) -> np.ndarray:
    lo = np.asarray(tp.min_force, F)
    hi = np.asarray(tp.max_force, F)
    xmin, xmax = nodes["xmin"], nodes["xmax"]
    leaf = (nodes["cl"] == 0) & (nodes["cr"] == 0) & reachable(nodes)
    ok = ((xmax < hi) & (xmax > lo)) | ((xmin < hi) & (xmin > lo))
    inside = leaf & np.all(ok, axis=1)
    ids = np.nonzero(inside)[0]
    return ids[np.argsort(-nodes["offset"][ids], kind="stable")]


def check_leaf_order(tp, tree) -> None:
    got = O.leaf_order(tp, tree)
    want = expected_leaf_order(tp, tree.nodes)
    assert np.array_equal(got, want), "leaf selection or order differs"


def seg_ranges(nodes, a, b):
    """(first particle, count) of each list segment."""
    lo = np.where(b >= 0, a, nodes["offset"][np.where(b >= 0, 0, a)])
    n = np.where(b >= 0, b, nodes["count"][np.where(b >= 0, 0, a)])
    return lo, n


def check_list_order(tree, lists) -> None:
    nodes = tree.nodes
    for k, tl in enumerate(lists.leaves):
        s0, s1 = lists.seg_ptr[k], lists.seg_ptr[k + 1]
        a, b = lists.seg_a[s0:s1], lists.seg_b[s0:s1]
        lo, n = seg_ranges(nodes, a, b)
        own = (nodes["offset"][tl], nodes["count"][tl])
        assert s1 > s0 and b[-1] >= 0 and (lo[-1], n[-1]) == own, \
            f"leaf {tl}: its own particles are not the last segment"
        if tl == 0:  # a root leaf walks itself first
            assert s1 - s0 == 2 and (lo[0], n[0]) == own
            continue
        lo, n = lo[:-1], n[:-1]
        keep = n > 0
        lo, hi = lo[keep], lo[keep] + n[keep]
        assert np.all((hi <= own[0]) | (lo >= own[0] + own[1])), \
            f"leaf {tl}: a segment overlaps the leaf itself"
        assert np.all(hi[1:] <= lo[:-1]), \
            f"leaf {tl}: segments are not in decreasing offset order"


def walk_decisions(tp, nodes, tl):
    """Per-node opening (looksBig), far-cut and L-inf closeness for target
    leaf tl, in float32 as the reference evaluates them."""
    tan = F(tp.tan_oa())
    fsr = F(tp.fsrrmax)
    xc, xmin, xmax = nodes["xc"], nodes["xmin"], nodes["xmax"]
    c = xc[tl]
    d = xc - c
    dist2 = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]
    s = xmax - xmin
    l2 = np.minimum(s[:, 0] * s[:, 0], np.minimum(s[:, 1] * s[:, 1], s[:, 2] * s[:, 2]))
    dtt2 = dist2 * tan * tan
    big = l2 > dtt2
    # the four space diagonals of the node's box, z fixed min -> max
    for i in (0, 1):
        for j in (0, 1):
            x1 = (xmin if i == 0 else xmax)[:, 0] - c[0]
            y1 = (xmin if j == 0 else xmax)[:, 1] - c[1]
            z1 = xmin[:, 2] - c[2]
            x2 = (xmax if i == 0 else xmin)[:, 0] - c[0]
            y2 = (xmax if j == 0 else xmin)[:, 1] - c[1]
            z2 = xmax[:, 2] - c[2]
            dx, dy, dz = x1 - x2, y1 - y2, z1 - z2
            big |= (dx * dx + dy * dy + dz * dz) > dtt2
    far = dist2 > fsr * fsr
    lo, hi = xmin[tl], xmax[tl]
    gap = np.where(hi < xmin, xmin - hi, np.where(lo > xmax, lo - xmax, F(0)))
    close = ~np.any(gap > fsr, axis=1)
    return big, far, close


def expected_list(tp, nodes, tl):
    """Segments (a, b) of leaf tl's list from a recursive walk."""
    T = max(1, int(tp.tdpts))
    cl, cr, cnt, off = nodes["cl"], nodes["cr"], nodes["count"], nodes["offset"]
    big, far, close = walk_decisions(tp, nodes, tl)
    segs = []

    def walk(v):
        if not big[v]:
            if far[v]:
                return
            segs.append((off[v], cnt[v]) if cnt[v] <= T else (v, -1))
            return
        if cl[v] == 0 and cr[v] == 0:
            segs.append((off[v], cnt[v]))
            return
        for ch in (cr[v], cl[v]):  # the right child's subtree comes first
            if ch > 0 and cnt[ch] > 0 and close[ch]:
                walk(int(ch))

    # ancestors of tl from the root, and the siblings of the path
    right, left = [], []
    v = 0
    while v != tl:
        a, b = cl[v], cr[v]
        on_left = a > 0 and off[a] <= off[tl] < off[a] + cnt[a]
        nxt, sib = (a, b) if on_left else (b, a)
        if sib > 0 and cnt[sib] > 0:
            (right if on_left else left).append(int(sib))
        v = int(nxt)
    roots = [0] if tl == 0 else right + left[::-1]
    for r in roots:
        walk(r)
    segs.append((off[tl], cnt[tl]))
    return segs


def check_list_walk(tp, tree, lists, max_leaves=None) -> None:
    nodes = tree.nodes
    ks = range(len(lists.leaves))
    if max_leaves is not None and len(lists.leaves) > max_leaves:
        ks = np.random.default_rng(0).choice(len(lists.leaves), max_leaves, replace=False)
    for k in ks:
        tl = int(lists.leaves[k])
        s0, s1 = lists.seg_ptr[k], lists.seg_ptr[k + 1]
        got = list(zip(lists.seg_a[s0:s1].tolist(), lists.seg_b[s0:s1].tolist()))
        want = [(int(a), int(b)) for a, b in expected_list(tp, nodes, tl)]
        assert got == want, f"leaf {tl}: list differs from the recursive walk"


def check_pptdr(tp, tree) -> None:
    nodes = tree.nodes
    r = reachable(nodes)
    xc, xmin, xmax = nodes["xc"][r], nodes["xmin"][r], nodes["xmax"][r]
    faces = np.concatenate([xmax - xc, xc - xmin], axis=1)
    want = F(tp.ppc) * faces.min(axis=1)
    got = nodes["tdr"][r]
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), "tdr differs"


def poly_literals():
    with open(os.path.join(HERE, "golden", "ref_poly_coeffs.json")) as f:
        d = json.load(f)
    as_float = lambda v: np.array([float(s) for s in v], np.float64).astype(F)
    return {k: as_float(v) for k, v in d.items() if isinstance(v, list)}


def check_law_constants() -> None:
    lit = poly_literals()
    assert np.array_equal(lit["step16_scalar"], lit["step16_qpx"])
    assert np.array_equal(lit["step16_scalar"], lit["forcelaw_poly5"][:6])
    for law, key, k in ((O.LAW_STEP16, "step16_scalar", 6), (O.LAW_SR_POLY5, "forcelaw_poly5", 6),
                        (O.LAW_SR_POLY6, "forcelaw_poly6", 7)):
        c = O.law_coeffs(law)
        assert c.shape[0] == k and np.array_equal(c.view(np.uint32), lit[key][:k].view(np.uint32)), \
            f"coefficients of law {law} differ from the reference's literals"


# ------------------------------------------------------------------ inputs
def lattice(nx, ny, nz, step=1.0, dup=1):
    """Every point of an nx x ny x nz lattice (dup copies each), unit masses:
    node boxes, centres and distances are small dyadic numbers, so the walk's
    comparisons meet exact ties."""
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1)
    p = (g.reshape(-1, 3) * step).astype(F)
    p = np.repeat(p, dup, axis=0)
    return p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy()


def tie_cases():
    """(name, TreeParams, x, y, z): inputs whose walk decisions tie."""
    out = []
    x, y, z = lattice(8, 8, 8)
    for tan, fsr in ((0.5, 2.0), (1.0, 2.0), (0.5, 3.0), (0.25, 4.0)):
        tp = O.TreeParams(np.zeros(3, F), np.full(3, 7, F), np.full(3, 1, F), np.full(3, 6, F),
                          fsrrmax=fsr, rsm=0.1, nd=8, tan=tan)
        out.append((f"lattice8 tan {tan} fsr {fsr}", tp, x, y, z))
    x, y, z = lattice(12, 6, 9)
    tp = O.TreeParams(np.zeros(3, F), np.array([11, 5, 8], F), np.full(3, 2, F), np.array([9, 4, 6], F),
                      fsrrmax=2.0, rsm=0.1, nd=4, tan=0.5)
    out.append(("lattice 12x6x9 nd 4", tp, x, y, z))
    x, y, z = lattice(8, 8, 8, step=0.5, dup=2)
    tp = O.TreeParams(np.zeros(3, F), np.full(3, 3.5, F), np.full(3, 0.5, F), np.full(3, 3, F),
                      fsrrmax=1.0, rsm=0.1, nd=16, tan=0.5)
    out.append(("half lattice, doubled points", tp, x, y, z))
    rng = np.random.default_rng(5)
    p = (rng.integers(0, 40, size=(6000, 3)) * 0.25).astype(F)
    tp = O.TreeParams(np.zeros(3, F), np.full(3, 9.75, F), np.full(3, 1, F), np.full(3, 8.5, F),
                      fsrrmax=1.5, rsm=0.1, nd=16, tan=0.5)
    out.append(("quarter grid random", tp, p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy()))
    # clumps of coincident particles become degenerate leaves whose box is a
    # point (every edge and diagonal 0); with tan = 0 every opening test
    # compares 0 with 0
    q = (rng.random((1500, 3)) * 8).astype(F)
    clumps = np.repeat((rng.random((40, 3)) * 8).astype(F), 5, axis=0)
    p = np.concatenate([q, clumps])
    tp = O.TreeParams(np.zeros(3, F), np.full(3, 8, F), np.full(3, 1, F), np.full(3, 7, F),
                      fsrrmax=1.5, rsm=0.1, nd=2, tan=0.0)
    out.append(("coincident clumps tan 0", tp, p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy()))
    return out


def generic_cases():
    out = []
    rng = np.random.default_rng(17)
    p = (rng.random((20000, 3)) * 16).astype(F)
    tp = O.TreeParams(np.zeros(3, F), np.full(3, 16, F), np.full(3, 2, F), np.full(3, 14, F),
                      fsrrmax=3.12, rsm=0.1, oa=0.1, nd=32)
    out.append(("uniform 20k", tp, p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy()))
    c = rng.random((8, 3)) * 12 + 2
    q = (c[rng.integers(0, 8, 12000)] + rng.normal(0, 0.3, (12000, 3))).clip(0, 16).astype(F)
    p = np.concatenate([q, (rng.random((4000, 3)) * 16).astype(F)])
    tp = O.TreeParams(np.zeros(3, F), np.full(3, 16, F), np.full(3, 2, F), np.full(3, 14, F),
                      fsrrmax=3.12, rsm=0.1, oa=0.5, nd=16)
    out.append(("clumps oa 0.5", tp, p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy()))
    tp12 = O.TreeParams(np.zeros(3, F), np.full(3, 16, F), np.full(3, 2, F), np.full(3, 14, F),
                        fsrrmax=3.12, rsm=0.1, oa=0.8, nd=32, tdpts=12)
    out.append(("clumps oa 0.8 TDPTS 12", tp12, p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy()))
    return out


def run_all(cases, max_leaves=None):
    """Every check on every case; returns the names of the checks that
    failed (empty when the oracle passes all)."""
    failed = []
    for name, tp, x, y, z in cases:
        tree = O.build_tree(tp, x, y, z)
        lists = O.build_lists(tp, tree)
        for label, fn in (("leaf_order", lambda: check_leaf_order(tp, tree)),
                          ("list_order", lambda: check_list_order(tree, lists)),
                          ("list_walk", lambda: check_list_walk(tp, tree, lists, max_leaves)),
                          ("pptdr", lambda: check_pptdr(tp, tree))):
            try:
                fn()
            except AssertionError as e:
                failed.append(f"{name}: {label}: {e}")
    try:
        check_law_constants()
    except AssertionError as e:
        failed.append(f"law constants: {e}")
    return failed
