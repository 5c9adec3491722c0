"""CPU tests pinning the oracle (oracle/hsr_oracle.c).

The reference's RCBForceTree.cxx cannot be compiled here (it needs MPI
headers and a generated CosmoDefinition.h), so the oracle is pinned by:
  * the reference's own BGQCM.c and ForceLaw.cxx compiled into oracle/_ref
    (live, where the reference tree is mounted) and the golden vectors they
    wrote into tests/golden (everywhere);
  * independent formulations of the partition (swap-chain destinations);
  * structural invariants of the tree;
  * the oa -> 0 limit of the lists, which must reproduce a brute-force
    direct sum, and disjoint coverage of every list.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def f32bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


# ------------------------------------------------------------------ cm
def test_cm_golden_vectors():
    d = np.load(os.path.join(GOLDEN, "ref_cm.npz"))
    pos, mass, offs, out = d["pos"], d["mass"], d["offsets"], d["out"]
    for k in range(offs.shape[0] - 1):
        a, b = offs[k], offs[k + 1]
        p = pos[a:b]
        xmin, xmax, xc = O.cm(p[:, 0], p[:, 1], p[:, 2], mass[a:b])
        got = np.concatenate([xmin, xmax, xc])
        assert np.array_equal(f32bits(got), f32bits(out[k])), f"case {k}"


def test_cm_live_reference_random():
    R = O.ref()
    if R is None:
        pytest.skip("oracle/_ref not built (reference tree not mounted)")
    rng = np.random.default_rng(5)
    for trial in range(200):
        n = int(rng.integers(1, 3000))
        scale = np.exp2(rng.integers(-30, 12, size=(n, 3)))
        p = (rng.random((n, 3)) * scale).astype(np.float32)
        w = np.ones(n, np.float32) if trial % 2 else (rng.random(n) + 0.5).astype(np.float32)
        x, y, z = (np.ascontiguousarray(p[:, k]) for k in range(3))
        ref = [np.zeros(3, np.float32) for _ in range(3)]
        R.hsrref_cm(n, x, y, z, w, *ref)
        got = O.cm(x, y, z, w)
        for g, r in zip(got, ref):
            assert np.array_equal(f32bits(g), f32bits(r))


# ------------------------------------------------------------ force laws
def test_sr_poly6_law_matches_reference_bitwise():
    d = np.load(os.path.join(GOLDEN, "ref_forcelaw.npz"))
    got = O.law_f32(O.LAW_SR_POLY6, d["r2"], rsm=float(d["rsm"]))
    assert np.array_equal(f32bits(got), f32bits(d["sr_poly6"]))
    assert float(d["rmax"]) == np.float32(3.116326355)


def test_p5_coefficients_fit_reference_grid_force():
    """The 5th-order polynomial of BGQStep16.c approximates the reference's
    fitted grid force fgor (FGridEvalFit, compiled from ForceLaw.cxx) over
    [0, rmax]; flipping or dropping any one coefficient breaks the fit."""
    d = np.load(os.path.join(GOLDEN, "ref_forcelaw.npz"))
    r2 = d["r2"].astype(np.float64)
    sel = r2 <= float(d["rmax"]) ** 2
    rsm = float(d["rsm"])
    # recover P5 from the oracle's Step16 law: f = (r2+rsm2)^-1.5 - P5(r2)
    x = r2[sel][r2[sel] > 0]
    f = O.law_f32(O.LAW_STEP16, x, rsm=rsm, fsrrmax=3.12).astype(np.float64)
    p5 = (x + np.float32(rsm) ** 2) ** -1.5 - f
    fg = d["fgor"].astype(np.float64)[sel][r2[sel] > 0]
    err = np.max(np.abs(p5 - fg))
    assert err < 0.03, err
    c = np.polynomial.polynomial.polyfit(x, p5, 5)
    for k in range(6):
        for mod in (-1.0, 0.0):
            cc = c.copy()
            cc[k] *= mod
            e = np.max(np.abs(np.polynomial.polynomial.polyval(x, cc) - fg))
            assert e > 0.1, (k, mod, e)


def test_step16_law_known_answers():
    rsm, rc = 0.1, 3.12
    a0 = np.float32(0.269327)
    f = O.law_f32(O.LAW_STEP16, [0.0, 1e-12, rc * rc, 10.0, 9.0], rsm=rsm, fsrrmax=rc)
    assert f[0] == 0.0                       # r2 > 0 guard (BGQStep16.c:182)
    assert abs(f[1] - (np.float32(rsm) ** -3 - a0)) < 1e-3 * np.float32(rsm) ** -3
    assert f[2] == 0.0 and f[3] == 0.0       # r2 < fsrrmax^2 strictly
    s = 9.0 + 0.01
    p = sum(c * 9.0 ** i for i, c in enumerate([0.269327, -0.0750978, 0.0114808,
                                                -0.00109313, 0.0000605491, -0.00000147177]))
    assert abs(f[4] - (s ** -1.5 - p)) < 1e-6


# --------------------------------------------------------------- partition
def test_partition_invariants():
    rng = np.random.default_rng(9)
    for trial in range(300):
        n = int(rng.integers(1, 400))
        key = rng.integers(0, 8, n).astype(np.float32)
        pv = float(rng.integers(0, 9))
        is_, out, ids = O.partition(key, pv)
        left = np.nonzero(key < pv)[0]
        assert is_ == left.shape[0]
        # left side: the keys < pv in their original order
        assert np.array_equal(ids[:is_], left)
        # right side: elements originally at q >= is stay put unless they
        # are left elements; a right element from q < is ends at the first
        # idx-iterate q -> idx[q] -> ... that is >= is
        dest = np.empty(n, np.int64)
        dest[ids] = np.arange(n)
        isleft = key < pv
        for q in range(n):
            if isleft[q]:
                continue
            if q >= is_:
                assert dest[q] == q
            else:
                j = q
                while j < is_:
                    j = left[j]
                assert dest[q] == j
        assert np.array_equal(np.sort(ids), np.arange(n))


# -------------------------------------------------------------------- tree
def uniform(n, box, seed):
    rng = np.random.default_rng(seed)
    p = (rng.random((n, 3)) * box).astype(np.float32)
    return p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy()


def small_params(box, nd, oa=0.1, margin=1.0):
    return O.TreeParams(np.zeros(3, np.float32), np.full(3, box, np.float32),
                        np.full(3, margin, np.float32), np.full(3, box - margin, np.float32),
                        nd=nd, oa=oa)


@pytest.mark.parametrize("n,nd,seed", [(5000, 16, 1), (20000, 64, 2), (3000, 1, 3)])
def test_tree_invariants(n, nd, seed):
    x, y, z = uniform(n, 10.0, seed)
    tp = small_params(10.0, nd)
    tree = O.build_tree(tp, x, y, z)
    t = tree.nodes
    R = O.ref()
    assert np.array_equal(np.sort(tree.id), np.arange(n))
    # allocation order: children of node k sit at 1 + 2A and 2 + 2A where A
    # counts the allocating nodes before k in preorder; here: children come
    # in adjacent pairs after their parent
    for k in range(t.shape[0]):
        cl, cr = t["cl"][k], t["cr"][k]
        if cl or cr:
            assert cl > k and cr == cl + 1
            assert t["count"][cl] + t["count"][cr] == t["count"][k]
            assert t["offset"][cl] == t["offset"][k]
            assert t["offset"][cr] == t["offset"][k] + t["count"][cl]
            # split on the longest edge, ties x > y > z (RCBForceTree.cxx:843-852)
            xl = (t["xmax"][k] - t["xmin"][k]).astype(np.float32)
            d = 0 if (xl[0] > xl[1] and xl[0] > xl[2]) else (1 if xl[1] > xl[2] else 2)
            coord = [tree.x, tree.y, tree.z][d]
            a, m_, b = t["offset"][k], t["offset"][cr], t["offset"][k] + t["count"][k]
            assert np.all(coord[a:m_] < t["xc"][k][d])
            assert np.all(coord[m_:b] >= t["xc"][k][d])
        elif t["count"][k] > 0:
            assert t["count"][k] <= nd or np.all(tree.x[t["offset"][k]:t["offset"][k] + t["count"][k]] == tree.x[t["offset"][k]])
            # leaf order is final at cm time, so the reference's cm over the
            # leaf's range reproduces the node bit for bit
            a, b = t["offset"][k], t["offset"][k] + t["count"][k]
            if R is not None:
                ref = [np.zeros(3, np.float32) for _ in range(3)]
                R.hsrref_cm(b - a, tree.x[a:b].copy(), tree.y[a:b].copy(), tree.z[a:b].copy(),
                            tree.mass[a:b].copy(), *ref)
                assert np.array_equal(f32bits(ref[0]), f32bits(t["xmin"][k]))
                assert np.array_equal(f32bits(ref[1]), f32bits(t["xmax"][k]))
                assert np.array_equal(f32bits(ref[2]), f32bits(t["xc"][k]))
            # tight bbox
            assert np.float32(tree.x[a:b].min()) == t["xmin"][k][0]
            assert np.float32(tree.z[a:b].max()) == t["xmax"][k][2]
            # TDPTS = 1 pseudo-particle mass = member mass (when count > 1)
            if t["count"][k] > 1:
                assert t["ppm"][k][0] == np.float32(t["count"][k])


def test_degenerate_nodes_stay_leaves():
    """Duplicate positions cannot be split: the node stays a leaf with more
    than nd particles and its two reserved child slots stay empty
    (RCBForceTree.cxx:727-729, 825-835)."""
    x = np.full(50, 2.5, np.float32)
    y = np.full(50, 3.5, np.float32)
    z = np.full(50, 4.5, np.float32)
    x2, y2, z2 = uniform(200, 10.0, 4)
    tp = small_params(10.0, 8)
    tree = O.build_tree(tp, np.concatenate([x, x2]), np.concatenate([y, y2]), np.concatenate([z, z2]))
    t = tree.nodes
    big = np.nonzero((t["count"] > 8) & (t["cl"] == 0) & (t["cr"] == 0))[0]
    assert big.size >= 1
    assert np.any(t["count"] == 0)  # orphan slots
    k = big[0]
    assert t["ppm"][k][0] == 0.0


# ------------------------------------------------------------------- lists
def coverage(tree, lists, k):
    """Particles represented by list k: direct entries and pseudo nodes."""
    n = tree.x.shape[0]
    cov = np.zeros(n, np.int64)
    a, b = lists.seg_ptr[k], lists.seg_ptr[k + 1]
    segs = list(zip(lists.seg_a[a:b], lists.seg_b[a:b]))
    for sa, sb in segs[:-1]:  # the last segment is the leaf itself
        if sb >= 0:
            cov[sa:sa + sb] += 1
        else:
            nd = tree.nodes[sa]
            cov[nd["offset"]:nd["offset"] + nd["count"]] += 1
    return cov


def test_lists_disjoint_coverage():
    x, y, z = uniform(20000, 16.0, 5)
    tp = small_params(16.0, 32)
    tree = O.build_tree(tp, x, y, z)
    lists = O.build_lists(tp, tree)
    for k in range(0, lists.leaves.shape[0], max(1, lists.leaves.shape[0] // 40)):
        cov = coverage(tree, lists, k)
        assert cov.max() <= 1
        nd = tree.nodes[lists.leaves[k]]
        assert np.all(cov[nd["offset"]:nd["offset"] + nd["count"]] == 0)
        # list length = direct entries + pseudo-particles + own particles
        a, b = lists.seg_ptr[k], lists.seg_ptr[k + 1]
        sb = lists.seg_b[a:b]
        assert lists.entries[k] == np.sum(np.where(sb >= 0, sb, 1))


def test_lists_oa0_reproduce_direct_sum():
    """With tan(oa) = 0 no node is ever accepted, so each list holds every
    particle within the L-infinity reach of the leaf: the tree force must
    equal the brute-force direct sum (FP64, different summation order)."""
    x, y, z = uniform(6000, 12.0, 6)
    m = np.ones_like(x)
    tp = small_params(12.0, 32, oa=0.0)
    tree = O.build_tree(tp, x, y, z)
    lists = O.build_lists(tp, tree)
    assert np.all(lists.seg_b >= 0)  # no pseudo-particles
    ax, ay, az = O.forces_f64(tp, tree, lists)
    targets = np.concatenate([np.arange(tree.nodes[l]["offset"], tree.nodes[l]["offset"] + tree.nodes[l]["count"])
                              for l in lists.leaves])
    dx, dy, dz = O.direct_f64(tree.x, tree.y, tree.z, tree.mass, targets)
    G = np.stack([ax[targets], ay[targets], az[targets]], 1)
    D = np.stack([dx, dy, dz], 1)
    assert np.max(np.abs(G - D)) <= 1e-9 * np.max(np.abs(D))


def test_tree_force_accuracy_vs_direct():
    """oa = 0.1 pseudo-particle approximation stays close to the direct sum."""
    x, y, z = uniform(8000, 12.0, 7)
    tp = small_params(12.0, 32, oa=0.1)
    tree = O.build_tree(tp, x, y, z)
    lists = O.build_lists(tp, tree)
    ax, ay, az = O.forces_f64(tp, tree, lists)
    targets = np.concatenate([np.arange(tree.nodes[l]["offset"], tree.nodes[l]["offset"] + tree.nodes[l]["count"])
                              for l in lists.leaves])
    dx, dy, dz = O.direct_f64(tree.x, tree.y, tree.z, tree.mass, targets)
    G = np.stack([ax[targets], ay[targets], az[targets]], 1)
    D = np.stack([dx, dy, dz], 1)
    rel = np.sqrt(np.sum((G - D) ** 2) / np.sum(D ** 2))
    assert rel < 2e-2, rel


def test_float_kernel_vs_fp64():
    x, y, z = uniform(8000, 12.0, 8)
    tp = small_params(12.0, 32)
    tree = O.build_tree(tp, x, y, z)
    lists = O.build_lists(tp, tree)
    f32 = np.stack(O.forces_f32(tp, tree, lists), 1)
    s16 = np.stack(O.forces_step16(tp, tree, lists), 1)
    f64 = np.stack(O.forces_f64(tp, tree, lists), 1)
    sel = np.any(f64 != 0, axis=1)
    rel = lambda a, b: np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2))
    assert rel(f32[sel], f64[sel]) < 1e-5
    assert rel(s16[sel], f64[sel]) < 1e-5


# ------------------------------------------------- quadrupole tree (T = 12)
def _icosahedron():
    """Cyclic permutations of (0, +-phi, +-1), normalised."""
    phi = (1 + 5 ** 0.5) / 2
    v = []
    for a in (-phi, phi):
        for b in (-1, 1):
            v += [(0, a, b), (a, b, 0), (b, 0, a)]
    v = np.array(v) / np.sqrt(1 + phi ** 2)
    return v


def test_quadrupole_design_is_the_icosahedron():
    """The 12 pseudo-particle directions are the icosahedron's vertices (a
    spherical 4-design): unit vectors, sum zero, second moment K/3 I."""
    P = O.pppts(12, 1.0, np.zeros(3, np.float32)).astype(np.float64)
    assert np.allclose(np.linalg.norm(P, axis=1), 1, atol=1e-6)
    ico = _icosahedron()
    d = np.linalg.norm(P[:, None, :] - ico[None, :, :], axis=2)
    assert np.all(d.min(1) < 1e-6) and len(set(d.argmin(1))) == 12
    assert np.allclose(P.sum(0), 0, atol=1e-6)
    assert np.allclose(P.T @ P, 4 * np.eye(3), atol=1e-5)
    # a 4-design: the fourth moment along any axis is K/5
    a = np.random.default_rng(0).normal(size=3)
    a /= np.linalg.norm(a)
    assert abs(np.sum((P @ a) ** 4) - 12 / 5) < 1e-5


def _moments(pos, m, c):
    d = pos.astype(np.float64) - c
    m = m.astype(np.float64)
    M0 = m.sum()
    M1 = (m[:, None] * d).sum(0)
    Q = (m[:, None, None] * (3 * d[:, :, None] * d[:, None, :]
                             - (d ** 2).sum(1)[:, None, None] * np.eye(3))).sum(0)
    return M0, M1, Q


@pytest.mark.parametrize("n,nd,seed", [(6000, 64, 31), (20000, 32, 32)])
def test_quadrupole_pseudo_particles_reproduce_moments(n, nd, seed):
    """Every node's 12 pseudo-particles carry its members' mass, dipole and
    traceless quadrupole about xc (the point of RCBQuadrupoleForceTree); a
    wrong design constant, weight (1/K, 3/K, 5/K, the 0.5) or sign fails."""
    rng = np.random.default_rng(seed)
    x, y, z = uniform(n, 10.0, seed)
    m = rng.uniform(0.5, 2.0, n).astype(np.float32)
    tp = small_params(10.0, nd)
    tp.tdpts = 12
    tree = O.build_tree(tp, x, y, z, m)
    t, ppm = tree.nodes, tree.ppm_t
    assert ppm.shape == (t.shape[0], 12)
    assert np.array_equal(t["ppm"][:, 0], ppm[:, 0])
    pos = np.stack([tree.x, tree.y, tree.z], 1)
    checked = 0
    worst = [0.0, 0.0, 0.0]
    for k in range(t.shape[0]):
        cnt = t["count"][k]
        if cnt <= 12:
            assert np.all(ppm[k] == 0)
            continue
        leaf = t["cl"][k] == 0 and t["cr"][k] == 0
        if leaf and cnt > nd:          # degenerate: no pseudo mass
            assert np.all(ppm[k] == 0)
            continue
        c = t["xc"][k].astype(np.float64)
        sl = slice(t["offset"][k], t["offset"][k] + cnt)
        M0, M1, Q = _moments(pos[sl], tree.mass[sl], c)
        tdr = float(t["tdr"][k])
        P = O.pppts(12, tdr, t["xc"][k])
        P0, P1, PQ = _moments(P, ppm[k], c)
        # float rounding scales with the weights' magnitude: members far
        # outside the tdr sphere carry weights ~ (r/tdr)^2 of either sign
        rho = np.sqrt(((pos[sl] - c) ** 2).sum(1)) / tdr
        W = float(np.sum(tree.mass[sl] * (1 + 3 * rho + 5 * rho ** 2)))
        worst[0] = max(worst[0], abs(P0 - M0) / W)
        worst[1] = max(worst[1], np.abs(P1 - M1).max() / (W * tdr))
        worst[2] = max(worst[2], np.abs(PQ - Q).max() / (W * tdr * tdr))
        checked += 1
    assert checked > 20
    # rounding-level agreement (a wrong weight or design point gives O(1))
    assert worst[0] < 2e-5 and worst[1] < 2e-5 and worst[2] < 2e-5, worst


def test_quadrupole_tree_force_beats_monopole():
    """At a wide opening angle the quadrupole tree's force is much closer to
    the direct sum than the monopole tree's, and its lists count 12 entries
    per pseudo-particle and cover every particle at most once."""
    x, y, z = uniform(8000, 12.0, 33)
    errs = {}
    for T in (1, 12):
        tp = small_params(12.0, 8, oa=0.8)
        tp.tdpts = T
        tree = O.build_tree(tp, x, y, z)
        lists = O.build_lists(tp, tree)
        sb = lists.seg_b
        assert np.any(sb < 0)
        for k in range(lists.leaves.shape[0]):
            a, b = lists.seg_ptr[k], lists.seg_ptr[k + 1]
            assert lists.entries[k] == np.sum(np.where(sb[a:b] >= 0, sb[a:b], T))
            assert np.all(tree.nodes["count"][lists.seg_a[a:b][sb[a:b] < 0]] > T)
        ax, ay, az = O.forces_f64(tp, tree, lists)
        targets = np.concatenate([np.arange(tree.nodes[l]["offset"], tree.nodes[l]["offset"] + tree.nodes[l]["count"])
                                  for l in lists.leaves])
        dx, dy, dz = O.direct_f64(tree.x, tree.y, tree.z, tree.mass, targets)
        G = np.stack([ax[targets], ay[targets], az[targets]], 1)
        D = np.stack([dx, dy, dz], 1)
        errs[T] = np.sqrt(np.sum((G - D) ** 2) / np.sum(D ** 2))
        f32 = np.stack(O.forces_step16(tp, tree, lists), 1)[targets]
        assert np.sqrt(np.sum((f32 - G) ** 2) / np.sum(G ** 2)) < 1e-5
    assert errs[12] < 0.5 * errs[1], errs


# -------------------------------------------------------------------- CIC
def test_cic_weights_partition_unity_and_grid_points():
    """Each particle's eight CIC weights are in [0, 1] and sum to one; a
    particle on a grid point deposits into its own cell alone; the total
    deposit is c = gpscal^3 per particle, the outside slot taking the corners
    beyond the grid."""
    rng = np.random.default_rng(61)
    ng = [7, 6, 5]
    Ng = 7 * 6 * 5
    for _ in range(50):
        p = (rng.random(3) * (np.array(ng) - 1)).astype(np.float32)
        rho = O.cic(p[:1], p[1:2], p[2:], ng)
        w = rho[:Ng]
        assert np.all(w >= 0) and np.all(w <= 1) and rho[Ng] == 0
        assert abs(w.sum() - 1.0) < 1e-6
        assert np.count_nonzero(w) <= 8
    q = np.array([3.0], np.float32), np.array([2.0], np.float32), np.array([4.0], np.float32)
    rho = O.cic(*q, ng, gpscal=2.0)
    assert rho[(3 * 6 + 2) * 5 + 4] == 8.0 and np.count_nonzero(rho) == 1
    n = 4000
    x, y, z = (rng.uniform(-1.5, g + 0.5, n).astype(np.float32) for g in ng)
    rho = O.cic(x, y, z, ng, gpscal=1.5)
    assert abs(rho.astype(np.float64).sum() - 1.5 ** 3 * n) < 1e-4 * n


def test_inverse_cic_is_the_deposit_transpose_and_exact_on_linear_fields():
    """Interpolating a grid G at a particle equals the dot product of G with
    that particle's deposit weights (a swapped corner in either direction
    fails), and a linear field interpolates to its value at the particle."""
    rng = np.random.default_rng(62)
    ng = [8, 9, 10]
    Ng = 8 * 9 * 10
    G = rng.normal(size=Ng + 1).astype(np.float32)
    G[Ng] = 123.0                  # the outside slot reads as 0
    for _ in range(40):
        p = [np.array([rng.uniform(-0.9, g - 0.1)], np.float32) for g in ng]
        w = O.cic(*p, ng)
        want = float(np.dot(w[:Ng].astype(np.float64), G[:Ng]))
        got = O.inverse_cic(*p, ng, G, 1.0, 1.0, np.zeros(1, np.float32))[0]
        assert abs(got - want) <= 1e-5 * (1 + abs(want))
    i, j, k = np.meshgrid(np.arange(8), np.arange(9), np.arange(10), indexing="ij")
    lin = np.append((0.5 * i - 1.25 * j + 2.0 * k + 3.0).astype(np.float32).ravel(), 0)
    n = 500
    x, y, z = (rng.uniform(0, g - 1, n).astype(np.float32) for g in ng)
    v0 = rng.normal(size=n).astype(np.float32)
    v = O.inverse_cic(x, y, z, ng, lin, 0.25, 2.0, v0)
    exact = 0.5 * x.astype(np.float64) - 1.25 * y + 2.0 * z + 3.0
    assert np.allclose((v - v0) / 0.5, exact, rtol=0, atol=2e-5 * 30)


def test_cell_sort_is_stable_counting_sort():
    rng = np.random.default_rng(10)
    x = (rng.random(5000) * 9.0 - 0.5).astype(np.float32)
    y = (rng.random(5000) * 8.0).astype(np.float32)
    z = (rng.random(5000) * 8.0).astype(np.float32)
    perm, inside = O.cell_sort_perm(x, y, z, [8, 8, 8])
    ix, iy, iz = np.floor(x).astype(int), np.floor(y).astype(int), np.floor(z).astype(int)
    ok = (ix >= 0) & (ix < 8)
    key = np.where(ok, (ix * 8 + iy) * 8 + iz, 512)
    order = np.argsort(key, kind="stable")
    assert np.array_equal(perm[order], np.arange(5000))
    assert inside == ok.sum()


def test_map1_drift_closed_forms():
    """Particles::map1 restated: with alpha = adot = pp = 1 the prefactor is
    exactly 1 and the drift is x + tau * v in float; elsewhere the step factor
    follows 1 / (alpha adot pp^(1 + 1/alpha)) * tau to float rounding (an
    exponent of 1/alpha or 1 + alpha, or a dropped adot, fails this), and two
    drifts of tau/2 agree with one of tau to rounding."""
    rng = np.random.default_rng(21)
    n = 4097
    x, y, z = (rng.uniform(-4, 60, n).astype(np.float32) for _ in range(3))
    vx, vy, vz = (rng.normal(0, 3, n).astype(np.float32) for _ in range(3))
    assert O.map1_factor(1.0, 0.37, 1.0, 1.0) == float(np.float32(0.37))
    X, Y, Z = O.drift(x, y, z, vx, vy, vz, O.map1_factor(1.0, 0.37, 1.0, 1.0))
    t = np.float32(0.37)
    assert np.array_equal(X, x + t * vx) and np.array_equal(Z, z + t * vz)
    for pp, tau, adot, alpha in [(0.5, 0.02, 1.3, 1.0), (0.13, 0.004, 4.2, 1.5),
                                 (0.9, 0.1, 0.7, 0.75), (2.0, 0.01, 0.2, 2.0)]:
        want = tau / (alpha * adot * pp ** (1.0 + 1.0 / alpha))
        got = O.map1_factor(pp, tau, adot, alpha)
        assert abs(got - want) <= 4e-7 * abs(want), (pp, tau, adot, alpha)
    k = O.map1_factor(0.5, 0.02, 1.3, 1.0)
    h = O.map1_factor(0.5, 0.01, 1.3, 1.0)
    X1, Y1, Z1 = O.drift(x, y, z, vx, vy, vz, k)
    X2, Y2, Z2 = O.drift(*O.drift(x, y, z, vx, vy, vz, h), vx, vy, vz, h)
    assert np.allclose(X1, X2, rtol=0, atol=4e-6 * 64)


def test_map2_coeff_closed_form():
    """map2's kick coefficient gpscal^3 / (4 pi) * fscal * tau * stepFraction
    to float rounding (a dropped or doubled factor fails)."""
    import math
    for g, fs, tau, sf in [(1.0, 1.0, 1.0, 1.0), (2.0, 1.7, 0.03, 0.2),
                           (0.5, 0.8, 0.11, 1.0 / 3), (1.25, 2.3, 0.004, 0.5)]:
        want = g ** 3 / (4 * math.pi) * fs * tau * sf
        got = O.map2_coeff(g, fs, tau, sf)
        assert abs(got - want) <= 3e-7 * want, (g, fs, tau, sf)


def test_resort_oracle_moves_all_arrays_and_keeps_order():
    """oracle.resort applies the cell permutation to all ten arrays; within a
    cell the input order is kept and the tail holds every particle outside
    the grid."""
    rng = np.random.default_rng(12)
    n, ng = 3000, [5, 6, 7]
    st = {a: rng.uniform(-1, 8, n).astype(np.float32) for a in O.POSVEL}
    st["id"] = np.arange(n, dtype=np.int64)
    st["mask"] = rng.integers(0, 60000, n).astype(np.uint16)
    out, inside = O.resort(st, ng)
    fx, fy, fz = (np.floor(out[a]) for a in ("x", "y", "z"))
    ok = (fx >= 0) & (fx < 5) & (fy >= 0) & (fy < 6) & (fz >= 0) & (fz < 7)
    assert ok[:inside].all() and not ok[inside:].any()
    cell = ((fx * 6 + fy) * 7 + fz)[:inside]
    assert np.all(np.diff(cell) >= 0)
    same = np.diff(cell) == 0
    assert np.all(np.diff(out["id"][:inside])[same] > 0)
    assert np.all(np.diff(out["id"][inside:]) > 0)
    for a in list(O.POSVEL) + ["mask"]:
        assert np.array_equal(out[a], st[a][out["id"]])


# ------------------------------------------------- drop-in host force laws
def test_dropin_forcelaw_host_matches_reference_bitwise():
    """hacc-coral_amd/host/ForceLaw.cxx (the drop-in's host ForceLaw
    hierarchy) reproduces the reference's ForceLaw.cxx values bit for bit."""
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "hacc-coral_amd", "bin", "forcelaw_values")
    if not os.path.exists(exe):
        subprocess.check_call(["make", "-s", "-C", os.path.join(root, "hacc-coral_amd"),
                               os.path.join(root, "hacc-coral_amd", "bin", "forcelaw_values")])
    d = np.load(os.path.join(GOLDEN, "ref_forcelaw.npz"))
    inp = "\n".join(repr(float(v)) for v in d["r2"]) + "\n"
    out = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True).stdout.split("\n")
    rows = [r.split() for r in out if r and not r.startswith("rmax")]
    got = np.array([[int(h, 16) for h in r] for r in rows], np.uint32)
    assert got.shape[0] == d["r2"].shape[0]
    for k, name in enumerate(["sr_poly6", "sr_fit", "fgor", "poly6", "interp1024"]):
        assert np.array_equal(got[:, k], f32bits(d[name])), name


@pytest.mark.parametrize("n", [2, 97, 1024, 4096])
def test_dropin_fgrid_table_matches_reference_bitwise(n):
    """The table the drop-in library gives the device for FGridEvalInterp
    (hsr_cxx_fgrid_table over FGrid::fgor_r2_interp, ForceLaw.cxx:54-67) is
    the reference's table bit for bit."""
    import ctypes
    from haccsr import _lib
    lib = ctypes.CDLL(_lib.CXX_LIB_PATH)
    lib.hsr_cxx_fgrid_table.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    lib.hsr_cxx_fgrid_table.restype = None
    r2 = np.zeros(n, np.float32)
    f = np.zeros(n, np.float32)
    lib.hsr_cxx_fgrid_table(n, r2.ctypes.data, f.ctypes.data)
    d = np.load(os.path.join(GOLDEN, "ref_forcelaw.npz"))
    assert np.array_equal(f32bits(r2), f32bits(d[f"table{n}_r2"]))
    assert np.array_equal(f32bits(f), f32bits(d[f"table{n}_f"]))


# ------------------------------------------------------------------ adjoint
def test_adjoint_f64_matches_finite_differences():
    """The FP64 adjoint (including the pseudo-particle centre-of-mass term)
    equals central finite differences of the FP64 objective, h = 1e-3,
    on random coordinates (topology frozen)."""
    # a wide opening angle makes accepted pseudo-particle nodes common
    x, y, z = uniform(4000, 10.0, 13)
    tp = small_params(10.0, 8, oa=0.7)
    tree = O.build_tree(tp, x, y, z)
    lists = O.build_lists(tp, tree)
    assert np.sum(lists.seg_b < 0) > 100, "case must contain pseudo-particles"
    rng = np.random.default_rng(3)
    lam = rng.normal(size=(tree.x.shape[0], 3))
    sel = np.zeros(tree.x.shape[0], bool)
    for l in lists.leaves:
        nd = tree.nodes[l]
        sel[nd["offset"]:nd["offset"] + nd["count"]] = True
    lam[~sel] = 0.0
    grad = O.adjoint_f64(tp, tree, lists, lam)
    pos = [tree.x.astype(np.float64), tree.y.astype(np.float64), tree.z.astype(np.float64)]
    h = 1e-3
    # coordinates: random, plus members of accepted pseudo-particle nodes
    pseudo_nodes = np.unique(lists.seg_a[lists.seg_b < 0])
    members = [tree.nodes[n_]["offset"] for n_ in pseudo_nodes[:16]]
    idx = list(rng.choice(tree.x.shape[0], 48, replace=False)) + members
    errs, scale = [], np.max(np.abs(grad))
    for k in idx:
        d = int(rng.integers(0, 3))
        pp = [a.copy() for a in pos]
        pm = [a.copy() for a in pos]
        pp[d][k] += h
        pm[d][k] -= h
        fd = (O.objective_f64(tp, tree, lists, lam, pp) - O.objective_f64(tp, tree, lists, lam, pm)) / (2 * h)
        errs.append(abs(fd - grad[k, d]) / max(abs(grad[k, d]), 1e-3 * scale))
    assert np.median(errs) < 1e-4 and np.max(errs) < 5e-2, (np.median(errs), np.max(errs))


def test_force_f64_at_matches_tree_force():
    """force_f64_at (double positions, frozen topology) at the tree's own
    float positions equals hsro_force_f64 for direct entries; pseudo-
    particles then sit at the FP64 centre of mass instead of the float xc,
    so the two differ only at that rounding level."""
    x, y, z = uniform(3000, 10.0, 21)
    tp = small_params(10.0, 16, oa=0.5)
    tree = O.build_tree(tp, x, y, z)
    lists = O.build_lists(tp, tree)
    ax, ay, az = O.forces_f64(tp, tree, lists)
    pos = (tree.x.astype(np.float64), tree.y.astype(np.float64), tree.z.astype(np.float64))
    a = O.force_f64_at(tp, tree.nodes, lists, pos, tree.mass.astype(np.float64))
    ref = np.stack([ax, ay, az], 1)
    assert np.sqrt(np.sum((a - ref) ** 2) / np.sum(ref ** 2)) < 1e-5


@pytest.mark.parametrize("scheme", ["kd", "dkd"])
def test_kick_drift_reverse_pass_matches_finite_differences(scheme):
    """The FP64 reverse pass over 5 kick-drift (or drift-kick-drift) sub-steps
    with periodic images (tests/kdref.py, the reference gradient of the GPU
    autograd test) equals central finite differences of the FP64 trajectory,
    h = 1e-3, with every sub-step's topology frozen."""
    import kdref as K
    from haccsr import ics
    box, o, dt = 10.0, 3.2, 0.01
    kick = dt / (4 * np.pi)
    x, y, z = ics.clustered(10, box, 1, 31)
    load = ics.add_ghosts(x, y, z, box, o)
    base = np.stack([x, y, z], 1).astype(np.float64)
    idx = load.alive_index
    shift = np.stack([load.x, load.y, load.z], 1).astype(np.float64) - base[idx]
    rng = np.random.default_rng(9)
    v0, wx, wv = (rng.normal(size=base.shape) for _ in range(3))
    tp = O.TreeParams(np.full(3, -1, np.float32), np.full(3, box + 2 * o + 1, np.float32),
                      np.full(3, o, np.float32), np.full(3, box + o, np.float32), nd=32, oa=0.5)
    topos = K.oracle_topos(tp, base, v0, idx, shift, dt, kick, 5, scheme)
    assert sum(int(np.sum(t.lists.seg_b < 0)) for t in topos) > 100
    gx, gv = K.gradient(tp, topos, base, v0, idx, shift, dt, kick, wx, wv, scheme)
    coords = list(zip(rng.integers(0, base.shape[0], 24).tolist(), rng.integers(0, 3, 24).tolist()))
    fd = K.fd_gradient(tp, topos, base, v0, idx, shift, dt, kick, wx, wv, coords, scheme=scheme)
    w = np.array([wx[i, a] for i, a in coords])
    ref = np.array([gx[i, a] for i, a in coords]) - w
    err = np.sqrt(np.sum((fd - w - ref) ** 2) / np.sum(ref ** 2))
    assert err < 1e-3, err
    # a dropped image contribution or a wrong sub-step order would fail this
    assert np.sqrt(np.mean(ref ** 2)) > 1e-2


def test_adjoint_f64_coincident_pairs_match_finite_differences():
    """Exactly coincident particles: the r2 > 0 guard zeroes their pair force
    (BGQStep16.c:180) but with rsm > 0 the pair term is smooth through d = 0,
    so the adjoint keeps the derivative m g(rsm^2); finite differences of the
    FP64 objective at members of a 200-fold duplicate pin it (a dropped
    coincident-pair term is off by ~1e5 there)."""
    rng = np.random.default_rng(4)
    p = (rng.random((4000, 3)) * 10.0).astype(np.float32)
    dup = np.full(200, 5.25, np.float32)
    x = np.concatenate([p[:, 0], dup])
    y = np.concatenate([p[:, 1], dup + 0.5])
    z = np.concatenate([p[:, 2], dup - 0.5])
    tp = small_params(10.0, 32, oa=0.5)
    tree = O.build_tree(tp, x, y, z)
    lists = O.build_lists(tp, tree)
    lam = np.random.default_rng(3).normal(size=(tree.x.shape[0], 3))
    grad = O.adjoint_f64(tp, tree, lists, lam)
    pos = [tree.x.astype(np.float64), tree.y.astype(np.float64), tree.z.astype(np.float64)]
    members = np.nonzero((tree.x == np.float32(5.25)) & (tree.y == np.float32(5.75)))[0]
    assert members.size == 200
    h = 1e-4
    for k in members[:6]:
        for d in range(3):
            pp = [a.copy() for a in pos]
            pm = [a.copy() for a in pos]
            pp[d][k] += h
            pm[d][k] -= h
            fd = (O.objective_f64(tp, tree, lists, lam, pp)
                  - O.objective_f64(tp, tree, lists, lam, pm)) / (2 * h)
            assert abs(fd - grad[k, d]) <= 1e-4 * max(abs(fd), 1.0), (k, d, fd, grad[k, d])
