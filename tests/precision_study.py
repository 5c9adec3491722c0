"""Precision study of the float leaf kernel (run by hand on a GPU; not a
pytest module).

    python tests/precision_study.py [--n-side 512] [--subset 1000000] [--nd 256]

On the 512^3 clustered load (2000 halos, seed 1003; one rank's share with a
3.2 ghost shell) it reports, as one JSON line:
  * float kernel vs the device FP64 evaluation (hsr_forces_f64) over every
    listed particle: relative RMS, max abs and max relative error;
  * float kernel and device FP64 vs the CPU FP64 oracle (hsro_force_f64,
    OpenMP) on a seeded subset of listed leaves holding >= --subset particles;
  * the momentum check: per-leaf FP64 sums sum_i m_i a_i of the float and of
    the FP64 forces, their largest difference relative to the leaf's
    sum_i |m_i a_i|, and the total over the force box.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
import types

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path[:0] = [ROOT, os.path.join(ROOT, "hacc-coral_amd")]

from haccsr import Context, ics  # noqa: E402
from oracle import oracle as O  # noqa: E402


def log(msg):
    print(f"[precision] {msg}", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n-side", type=int, default=512)
    ap.add_argument("--subset", type=int, default=1_000_000)
    ap.add_argument("--nd", type=int, default=256)
    ap.add_argument("--seed", type=int, default=1003)
    a = ap.parse_args()
    t0 = time.perf_counter()
    load = ics.config_load("C", n_side=a.n_side, log=log)
    log(f"load: {load.n} particles ({load.n_alive} alive) in {time.perf_counter() - t0:.0f} s")
    lo, hi = load.tree_box
    flo, fhi = load.force_box
    n = load.n
    x, y, z = load.x.copy(), load.y.copy(), load.z.copy()
    m = np.ones(n, np.float32)
    ctx = Context(nd=a.nd)
    ctx.build(x, y, z, m, min_loc=lo, max_loc=hi)   # host arrays come back in tree order
    ctx.lists(flo, fhi)
    st = ctx.stats()
    log(f"tree: {st.nodes} nodes, {st.listed_leaves} listed leaves, {st.pairs:.3e} pairs")
    fx, fy, fz = ctx.forces()
    f32 = np.stack([fx, fy, fz], 1).astype(np.float64)
    log("float forces done")
    t1 = time.perf_counter()
    ax, ay, az, mom64 = ctx.forces_f64(with_momentum=True)
    f64 = np.stack([ax, ay, az], 1)
    log(f"device FP64 forces done in {time.perf_counter() - t1:.1f} s")
    mom32 = ctx.leaf_momentum(fx, fy, fz)

    leaves, seg_ptr, sa, sb, ent = ctx.get_lists()
    nodes = ctx.nodes().view(O.NODE_DTYPE)
    listed = np.zeros(n, bool)
    lnode = nodes[leaves]
    for off, cnt in zip(lnode["offset"], lnode["count"]):
        listed[off:off + cnt] = True
    d = f32[listed] - f64[listed]
    norm = np.linalg.norm(f64[listed], axis=1)
    dn = np.linalg.norm(d, axis=1)
    big = norm > 1e-3 * np.sqrt(np.mean(norm ** 2))
    full = {"particles": int(listed.sum()),
            "rel_rms": float(np.sqrt(np.sum(d ** 2) / np.sum(f64[listed] ** 2))),
            "max_abs": float(dn.max()),
            "max_rel": float((dn[big] / norm[big]).max()),
            "rms_force": float(np.sqrt(np.mean(norm ** 2)))}
    log(f"float vs device FP64: {full}")

    # CPU FP64 oracle on a seeded subset of listed leaves
    rng = np.random.default_rng(a.seed)
    order = rng.permutation(leaves.shape[0])
    cum = np.cumsum(lnode["count"][order])
    take = np.sort(order[: int(np.searchsorted(cum, a.subset)) + 1])
    sub_ptr = np.zeros(take.shape[0] + 1, np.int64)
    parts_a, parts_b = [], []
    for k, li in enumerate(take):
        parts_a.append(sa[seg_ptr[li]:seg_ptr[li + 1]])
        parts_b.append(sb[seg_ptr[li]:seg_ptr[li + 1]])
        sub_ptr[k + 1] = sub_ptr[k] + (seg_ptr[li + 1] - seg_ptr[li])
    sub = O.Lists(leaves[take].astype(np.int64), sub_ptr, np.concatenate(parts_a),
                  np.concatenate(parts_b), ent[take])
    tp = O.TreeParams(lo, hi, flo, fhi, nd=a.nd)
    tree = types.SimpleNamespace(nodes=nodes, x=x, y=y, z=z, mass=m)
    t2 = time.perf_counter()
    cx, cy, cz, cmom = O.forces_f64(tp, tree, sub, with_momentum=True)
    log(f"CPU FP64 oracle on {int(lnode['count'][take].sum())} particles in {time.perf_counter() - t2:.0f} s")
    sel = np.zeros(n, bool)
    for off, cnt in zip(nodes[sub.leaves]["offset"], nodes[sub.leaves]["count"]):
        sel[off:off + cnt] = True
    c64 = np.stack([cx, cy, cz], 1)[sel]
    e_dev = np.sqrt(np.sum((f64[sel] - c64) ** 2) / np.sum(c64 ** 2))
    d32 = f32[sel] - c64
    n64 = np.linalg.norm(c64, axis=1)
    big = n64 > 1e-3 * np.sqrt(np.mean(n64 ** 2))
    subset = {"particles": int(sel.sum()), "leaves": int(take.shape[0]),
              "device_fp64_vs_cpu_fp64_rel_rms": float(e_dev),
              "float_rel_rms": float(np.sqrt(np.sum(d32 ** 2) / np.sum(c64 ** 2))),
              "float_max_abs": float(np.linalg.norm(d32, axis=1).max()),
              "float_max_rel": float((np.linalg.norm(d32, axis=1)[big] / n64[big]).max()),
              "leaf_momentum_device_vs_cpu_max_abs": float(np.abs(mom64[take] - cmom).max())}
    log(f"subset: {subset}")

    # momentum check
    absmom = np.zeros((leaves.shape[0],))
    m64 = m.astype(np.float64)
    for k, (off, cnt) in enumerate(zip(lnode["offset"], lnode["count"])):
        absmom[k] = np.sum(m64[off:off + cnt] * np.linalg.norm(f64[off:off + cnt], axis=1))
    dm = np.linalg.norm(mom32 - mom64, axis=1)
    ok = absmom > 0
    momentum = {"leaf_max_rel_diff_float_vs_fp64": float((dm[ok] / absmom[ok]).max()),
                "leaf_median_rel_diff_float_vs_fp64": float(np.median(dm[ok] / absmom[ok])),
                "total_fp64_over_sum_abs": float(np.linalg.norm(mom64.sum(0)) / absmom.sum()),
                "total_float_over_sum_abs": float(np.linalg.norm(mom32.sum(0)) / absmom.sum())}
    log(f"momentum: {momentum}")
    print(json.dumps({"workload": f"{a.n_side}^3 clustered (2000 halos scaled by volume), "
                                  f"+{n - load.n_alive} ghosts, nd {a.nd}, eps 0.1, fsrrmax 3.12, oa 0.1",
                      "pairs": int(st.pairs), "full": full, "subset": subset,
                      "momentum": momentum}), flush=True)


if __name__ == "__main__":
    main()
