"""Precision study at one rank's production share: the 512^3 clustered load
(2000 halos, seed 1003, 3.2 ghost shell: 143M particles), leaf size 256.

  * float leaf kernel vs the device FP64 evaluation of the same lists
    (hsr_forces_f64) over every listed particle: relative RMS, largest
    absolute and relative error;
  * float kernel and device FP64 vs the CPU FP64 oracle (hsro_force_f64) on a
    seeded subset of listed leaves holding at least 1M particles (the device
    FP64 must equal the oracle to the last few bits: different summation
    order, same terms);
  * the momentum check: per-leaf FP64 sums of m_i a_i for the float and for
    the FP64 forces (hsr_leaf_momentum, hsr_forces_f64), their largest
    difference relative to the leaf's sum of |m_i a_i|.
The figures are printed; the bounds are the force tolerance (1e-5 relative
RMS) and, for the per-leaf momentum, 1e-4 of the leaf's sum of |m_i a_i|."""
import time
import types

import numpy as np
import pytest

from haccsr import Context, ics
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ND = 256
SUBSET = 1_000_000


def say(msg):
    print(msg, flush=True)


@pytest.mark.timeout(1200)
def test_config_c_precision_study():
    t0 = time.perf_counter()
    say("case config C 512^3 clustered (2000 halos), nd=256: load")
    load = ics.config_load("C", log=say)
    say(f"  {load.n} particles ({load.n_alive} alive) in {time.perf_counter() - t0:.0f} s")
    lo, hi = load.tree_box
    flo, fhi = load.force_box
    n = load.n
    x, y, z = load.x, load.y, load.z          # permuted into tree order by the build
    m = np.ones(n, np.float32)
    del load
    ctx = Context(nd=ND)
    ctx.build(x, y, z, m, min_loc=lo, max_loc=hi)
    ctx.lists(flo, fhi)
    st = ctx.stats()
    say(f"  tree: {st.nodes} nodes, depth {st.max_depth}, {st.listed_leaves} listed leaves, "
        f"{st.pairs:.3e} pairs")
    fx, fy, fz = ctx.forces()
    t1 = time.perf_counter()
    ax, ay, az, mom64 = ctx.forces_f64(with_momentum=True)
    say(f"  device FP64 forces in {time.perf_counter() - t1:.1f} s")
    mom32 = ctx.leaf_momentum(fx, fy, fz)
    f32 = np.stack([fx, fy, fz], 1).astype(np.float64)
    f64 = np.stack([ax, ay, az], 1)
    del fx, fy, fz, ax, ay, az

    leaves, seg_ptr, sa, sb, ent = ctx.get_lists()
    nodes = ctx.nodes().view(O.NODE_DTYPE)
    lnode = nodes[leaves]
    # particle ranges of the listed leaves (disjoint)
    listed = np.zeros(n + 1, np.int32)
    np.add.at(listed, lnode["offset"], 1)
    np.add.at(listed, lnode["offset"] + lnode["count"], -1)
    listed = np.cumsum(listed[:n]) > 0
    d = f32[listed] - f64[listed]
    norm = np.linalg.norm(f64[listed], axis=1)
    dn = np.linalg.norm(d, axis=1)
    big = norm > 1e-3 * np.sqrt(np.mean(norm ** 2))
    rel_rms = float(np.sqrt(np.sum(d ** 2) / np.sum(f64[listed] ** 2)))
    say(f"  float vs device FP64 over {int(listed.sum())} listed particles: rel-RMS {rel_rms:.3e}, "
        f"max abs {dn.max():.3e} (rms |F| {np.sqrt(np.mean(norm ** 2)):.3e}), "
        f"max rel {(dn[big] / norm[big]).max():.3e} where |F| > 1e-3 rms")
    assert rel_rms <= 1e-5
    del d, norm, dn, big

    # CPU FP64 oracle on a seeded subset of listed leaves
    rng = np.random.default_rng(1003)
    order = rng.permutation(leaves.shape[0])
    cum = np.cumsum(lnode["count"][order])
    take = np.sort(order[: int(np.searchsorted(cum, SUBSET)) + 1])
    sub_ptr = np.zeros(take.shape[0] + 1, np.int64)
    sub_ptr[1:] = np.cumsum(seg_ptr[take + 1] - seg_ptr[take])
    sub_a = np.concatenate([sa[seg_ptr[li]:seg_ptr[li + 1]] for li in take])
    sub_b = np.concatenate([sb[seg_ptr[li]:seg_ptr[li + 1]] for li in take])
    sub = O.Lists(leaves[take].astype(np.int64), sub_ptr, sub_a, sub_b, ent[take])
    tp = O.TreeParams(np.asarray(lo, np.float32), np.asarray(hi, np.float32),
                      np.asarray(flo, np.float32), np.asarray(fhi, np.float32), nd=ND)
    tree = types.SimpleNamespace(nodes=nodes, x=x, y=y, z=z, mass=m)
    t2 = time.perf_counter()
    cx, cy, cz, cmom = O.forces_f64(tp, tree, sub, with_momentum=True)
    sel = np.zeros(n, bool)
    for off, cnt in zip(nodes[sub.leaves]["offset"], nodes[sub.leaves]["count"]):
        sel[off:off + cnt] = True
    c64 = np.stack([cx, cy, cz], 1)[sel]
    e_dev = float(np.sqrt(np.sum((f64[sel] - c64) ** 2) / np.sum(c64 ** 2)))
    d32 = f32[sel] - c64
    e32 = float(np.sqrt(np.sum(d32 ** 2) / np.sum(c64 ** 2)))
    mom_dev = float(np.abs(mom64[take] - cmom).max() / max(np.abs(cmom).max(), 1e-300))
    say(f"  CPU FP64 oracle on {int(sel.sum())} particles ({take.shape[0]} leaves) in "
        f"{time.perf_counter() - t2:.0f} s: device FP64 vs oracle rel-RMS {e_dev:.3e}; "
        f"float vs oracle rel-RMS {e32:.3e}; leaf momentum device vs oracle {mom_dev:.3e}")
    assert sel.sum() >= SUBSET
    assert e_dev <= 1e-12 and mom_dev <= 1e-9
    assert e32 <= 1e-5

    # momentum check: per-leaf sums of m_i a_i, float vs FP64
    starts = lnode["offset"]
    counts = lnode["count"]
    absf = m.astype(np.float64) * np.linalg.norm(f64, axis=1)
    csum = np.concatenate([[0.0], np.cumsum(absf)])
    absmom = csum[starts + counts] - csum[starts]
    dm = np.linalg.norm(mom32 - mom64, axis=1)
    ok = absmom > 0
    rel = dm[ok] / absmom[ok]
    tot64 = float(np.linalg.norm(mom64.sum(0)) / absmom.sum())
    tot32 = float(np.linalg.norm(mom32.sum(0)) / absmom.sum())
    say(f"  momentum: per-leaf |sum m a (float) - sum m a (FP64)| / sum |m a|: max {rel.max():.3e}, "
        f"median {np.median(rel):.3e}; total |sum m a| / sum |m a|: FP64 {tot64:.3e}, float {tot32:.3e}")
    assert rel.max() <= 1e-4
    ctx.close()
