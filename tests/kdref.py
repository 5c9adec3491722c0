"""FP64 kick-drift with frozen per-substep topology: the reference trajectory,
its reverse-mode gradient and finite differences (test infrastructure).

One substep is v' = v + c F(x), x' = x + dt v' ("kd"), or x += dt/2 v,
v += c F(x), x += dt/2 v ("dkd", the half stream / kick / half stream of
Particles::subCycle, Particles.cxx:1176-1201), with c the kick coefficient of
map2 (Particles.cxx:1231-1234).  F acts on the
rank's full array (alive particles followed by periodic images, entry k =
alive particle index[k] shifted by shift[k]); only the alive rows of F are
used.  A topology is (perm, nodes, lists): the tree permutation (tree
position -> array index), the node table and the interaction lists that the
forward pass used at that substep.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from oracle import oracle as O


@dataclasses.dataclass
class Topo:
    perm: np.ndarray
    nodes: np.ndarray
    lists: "O.Lists"


def topo_from_oracle(tp, pos_all: np.ndarray) -> Topo:
    """Tree + lists of the oracle at float32(pos_all)."""
    p = pos_all.astype(np.float32)
    tree = O.build_tree(tp, p[:, 0], p[:, 1], p[:, 2], np.ones(p.shape[0], np.float32),
                        ids=np.arange(p.shape[0], dtype=np.int64))
    return Topo(tree.id.copy(), tree.nodes, O.build_lists(tp, tree))


def topo_from_record(rec) -> Topo:
    """Topology recorded by haccsr.autograd.ForceSolver(record=True)."""
    lists = O.Lists(rec.leaves, rec.seg_ptr, rec.seg_a, rec.seg_b, rec.entries)
    return Topo(rec.perm, rec.nodes.view(O.NODE_DTYPE), lists)


def expand(x, index, shift):
    return x[index] + shift


def force_all(tp, topo: Topo, pos_all: np.ndarray) -> np.ndarray:
    """FP64 F at pos_all (array order), frozen topology."""
    s = pos_all[topo.perm]
    m = np.ones(s.shape[0])
    a = O.force_f64_at(tp, topo.nodes, topo.lists, (s[:, 0], s[:, 1], s[:, 2]), m)
    out = np.zeros_like(pos_all)
    out[topo.perm] = a
    return out


def vjp_all(tp, topo: Topo, pos_all: np.ndarray, lam_all: np.ndarray) -> np.ndarray:
    """FP64 d(sum lam . F)/dx at pos_all (array order), frozen topology."""
    s = pos_all[topo.perm]
    m = np.ones(s.shape[0])
    g = O.adjoint_f64_at(tp, topo.nodes, topo.lists, (s[:, 0], s[:, 1], s[:, 2]), m,
                         lam_all[topo.perm])
    out = np.zeros_like(pos_all)
    out[topo.perm] = g
    return out


def trajectory(tp, topos, x0, v0, index, shift, dt, kick, scheme="kd"):
    """Positions at which each substep's force is evaluated and the final
    (x, v).  "kd": v += kick F(x), x += dt v; "dkd": x += dt/2 v,
    v += kick F(x), x += dt/2 v."""
    x, v = x0.astype(np.float64).copy(), v0.astype(np.float64).copy()
    n = x.shape[0]
    xs = []
    for topo in topos:
        if scheme == "dkd":
            x = x + 0.5 * dt * v
        xs.append(x.copy())
        F = force_all(tp, topo, expand(x, index, shift))[:n]
        v = v + kick * F
        x = x + (0.5 * dt if scheme == "dkd" else dt) * v
    return xs, x, v


def oracle_topos(tp, x0, v0, index, shift, dt, kick, nsub, scheme="kd"):
    """Topologies the oracle builds along its own FP64 trajectory."""
    x, v = x0.astype(np.float64).copy(), v0.astype(np.float64).copy()
    n = x.shape[0]
    topos = []
    for _ in range(nsub):
        if scheme == "dkd":
            x = x + 0.5 * dt * v
        pos = expand(x, index, shift)
        topo = topo_from_oracle(tp, pos)
        topos.append(topo)
        v = v + kick * force_all(tp, topo, pos)[:n]
        x = x + (0.5 * dt if scheme == "dkd" else dt) * v
    return topos


def objective(tp, topos, x0, v0, index, shift, dt, kick, wx, wv, scheme="kd"):
    _, x, v = trajectory(tp, topos, x0, v0, index, shift, dt, kick, scheme)
    return float(np.sum(wx * x) + np.sum(wv * v))


def gradient(tp, topos, x0, v0, index, shift, dt, kick, wx, wv, scheme="kd"):
    """Reverse pass: dJ/dx0, dJ/dv0 for J = sum wx . x_N + wv . v_N."""
    xs, _, _ = trajectory(tp, topos, x0, v0, index, shift, dt, kick, scheme)
    n = x0.shape[0]
    gx, gv = wx.astype(np.float64).copy(), wv.astype(np.float64).copy()
    h = 0.5 * dt if scheme == "dkd" else dt
    for topo, x in zip(reversed(topos), reversed(xs)):
        gv = gv + h * gx                     # last drift
        lam = np.zeros((index.shape[0], 3))
        lam[:n] = kick * gv
        g_all = vjp_all(tp, topo, expand(x, index, shift), lam)
        # images carry their alive particle's gradient back to it
        gx = gx + np.stack([np.bincount(index, weights=g_all[:, k], minlength=n)
                            for k in range(3)], 1)
        if scheme == "dkd":
            gv = gv + h * gx                 # first half drift
    return gx, gv


def fd_gradient(tp, topos, x0, v0, index, shift, dt, kick, wx, wv, coords, h=1e-3,
                scheme="kd"):
    """Central differences of the FP64 objective w.r.t. x0 at the given
    (particle, axis) coordinates."""
    out = np.zeros(len(coords))
    for k, (i, a) in enumerate(coords):
        xp = x0.astype(np.float64).copy()
        xm = xp.copy()
        xp[i, a] += h
        xm[i, a] -= h
        jp = objective(tp, topos, xp, v0, index, shift, dt, kick, wx, wv, scheme)
        jm = objective(tp, topos, xm, v0, index, shift, dt, kick, wx, wv, scheme)
        out[k] = (jp - jm) / (2 * h)
    return out
