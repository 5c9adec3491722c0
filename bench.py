#!/usr/bin/env python3
"""Benchmark of the MI355X short-range force path (one GPU).

One step = one pass of the hot path over the resident 256^3 clustered load:
restore the unsorted particle arrays (device copy), RCB tree build (sorting
all ten arrays into leaf order), interaction lists of the leaves inside the
force box, and the leaf kernel applying the short-range kick.  The metric is
list pairs per second (pairs = sum over listed leaves of n_leaf * L_leaf, the
reference's own unit of work: RCBForceTree.cxx:1143, 604-605).

Prints one JSON line (see the repository's benchmark contract), with the
leaf kernel's roofline against the FP32 vector peak and the CPU restatement
of the reference kernel timed on the host as the baseline.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "hacc-coral_amd"))

FLOPS_PER_PAIR = 30          # d 3, r2 5, +eps2 1, rsqrt 1, cube 2, P5 10, sub 1, *m 1, acc 6
PEAK_FP32_TFLOPS = 157.3     # MI355X FP32 vector peak (packed); non-packed 78.6
PEAK_HBM_GBPS = 8000.0       # MI355X HBM3E
# HBM bytes per k_force launch measured with rocprofv3 (FETCH_SIZE and WRITE_SIZE,
# scaled by what they read for a plain device copy of known size), keyed by the
# workload (config, n_side, leaf size, TDPTS), the kernel shape and the first 16
# hex digits of the SHA-256 of the leaf kernel's source (csrc/force.hip): a
# figure is reported only for the build it was measured on.
MEASURED_TRAFFIC = {
    # profiles/k_force_256c_traffic.md
    ("B", 256, 512, 1, (2, 2, 128), "ba11f51afafb0511"): 144144650880,
    ("B", 256, 640, 1, (2, 2, 128), "ba11f51afafb0511"): 147090420416,
}


def _kernel_src_hash():
    import hashlib
    path = os.path.join(ROOT, "hacc-coral_amd", "csrc", "force.hip")
    try:
        return hashlib.sha256(open(path, "rb").read()).hexdigest()[:16]
    except OSError:
        return None


KERNEL_SRC = _kernel_src_hash()
# Algorithmic bytes of the tree build: 48 B per particle per level it is split at
# (cm reads x, y, z, m; partition reads and writes key + index) plus one 84 B
# gather of the ten payload arrays.
BUILD_BYTES_PER_LEVEL, BUILD_BYTES_GATHER = 48, 84


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="B",
                    help="B (256^3 clustered), A (64^3), D (128^3), E (ensemble of 64 boxes)")
    ap.add_argument("--n-side", type=int, default=None, help="override the load's n^(1/3)")
    ap.add_argument("--nd", type=int, default=0, help="leaf size (0 = autotuned)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--adjoint-reps", type=int, default=3,
                    help="timed hsr_adjoint calls on the last step's tree (0 = skip)")
    ap.add_argument("--adjoint-mode", type=int, default=0,
                    help="hsr_tuning.adjoint_mode: 0 automatic, 1 fused, 2 two-pass")
    ap.add_argument("--tdpts", type=int, default=1, choices=(1, 12),
                    help="pseudo-particles per node: 1 (monopole, the bench line) or 12 "
                         "(quadrupole; no adjoint leg)")
    ap.add_argument("--subcycle-nsub", type=int, default=5,
                    help="sub-steps of the device-resident sub-cycle leg (0: skip it)")
    ap.add_argument("--pm-reps", type=int, default=3,
                    help="timed CIC deposit/interpolation pairs of the PM leg (0: skip it)")
    ap.add_argument("--force-pass", type=int, default=0, choices=(0, 32, 64, 128, 256),
                    help="hsr_tuning.force_pass: targets per leaf-kernel pass (0: default)")
    ap.add_argument("--i-per-lane", type=int, default=0,
                    help="leaf-kernel shape: i-particles per lane (0: autotuned with the "
                         "unroll and tile)")
    ap.add_argument("--unroll", type=int, default=0, help="leaf-kernel shape: j entries per step")
    ap.add_argument("--tile", type=int, default=0, help="leaf-kernel shape: entries per LDS tile")
    ap.add_argument("--refresh-reps", type=int, default=3,
                    help="timed one-rank overload refreshes of the refresh leg (0: skip it)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run one plain device-to-device copy of BYTES (the "
                         "reference kernel for scaling FETCH_SIZE/WRITE_SIZE counters)")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def workload_name(cfg, load, nd):
    ns = round(load.n_alive ** (1 / 3))
    halos = {"B": 500 * (ns / 256) ** 3, "C": 2000 * (ns / 512) ** 3, "D": 500 * (ns / 256) ** 3}
    if cfg == "A":
        name = f"{ns}^3 Plummer mix (70% uniform, 30% in 20 clumps)"
    elif cfg in halos:
        name = (f"{ns}^3 clustered (GRF P(k)~k^-2, Zel'dovich rms 2, "
                f"{max(1, round(halos[cfg]))} NFW halos)")
    else:
        name = cfg
    return (f"{name}, box {load.box:g}, +{load.n - load.n_alive} ghosts "
            f"(shell {load.overload:g}), eps 0.1, fsrrmax 3.12, oa 0.1, nd {nd}")


def ensemble_name(ens, n, nd):
    sides = ens["sides"]
    return (f"ensemble of {ens['nboxes']} clustered boxes (n_side log-uniform in [32, 128]: "
            f"{min(sides)}..{max(sides)}; {ens['alive']} alive, {n} with ghosts), eps 0.1, "
            f"fsrrmax 3.12, oa 0.1, nd {nd}; one build, one list pass, one graph-replayed "
            f"leaf launch over all boxes")


def host_cpu():
    """lscpu's model name and physical core count of this host."""
    import subprocess
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            info[k.strip()] = v.strip()
    except Exception:
        pass
    try:
        phys = int(info.get("Core(s) per socket", "0")) * int(info.get("Socket(s)", "0"))
    except ValueError:
        phys = 0
    return {"cpu_model": info.get("Model name", "unknown"), "physical_cores": phys or None,
            "logical_cpus": os.cpu_count()}


def cpu_baseline(load, nd):
    """The oracle (CPU restatement of the reference's tree build, lists and
    float kernel; OpenMP over leaves for lists and kernel) on the host:
    the tree build of the whole load once, then lists + kernel on a random
    sample of the same tree's listed leaves, median of 5 after a warm-up, on
    all the threads this process may use and on one thread.  The sample is
    8 leaves per thread (at least 64) for all threads and 8 leaves for one
    thread: about 20-30 s of host time at 256^3."""
    from oracle import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    lo, hi = load.tree_box
    flo, fhi = load.force_box
    tp = O.TreeParams(lo, hi, flo, fhi, nd=nd)
    t0 = time.perf_counter()
    tree = O.build_tree(tp, load.x, load.y, load.z)
    t_tree = time.perf_counter() - t0
    leaves = O.leaf_order(tp, tree)
    total_pairs = O.build_lists(tp, tree, leaves).pairs(tree)
    rng = np.random.default_rng(7)
    order = rng.permutation(leaves.shape[0])

    def run(sub):
        t = time.perf_counter()
        L = O.build_lists(tp, tree, sub)
        O.forces_f32(tp, tree, L)
        return time.perf_counter() - t, L.pairs(tree)

    def median5(sub):
        run(sub)  # warm-up
        ts = sorted(run(sub)[0] for _ in range(5))
        return ts[2]

    # enough leaves that dynamic scheduling balances the threads (list
    # lengths spread by 100x between voids and halo cores), and for one
    # thread the same number of leaves per thread
    O.set_threads(threads)
    k = min(leaves.shape[0], max(64, 8 * threads))
    sub = np.sort(leaves[order[:k]])
    t_all = median5(sub)
    pairs = run(sub)[1]
    k1 = max(1, k // max(threads, 1))
    sub1 = np.sort(leaves[order[:k1]])
    pairs1 = run(sub1)[1]
    O.set_threads(1)
    t_one = median5(sub1)
    O.set_threads(threads)
    rate = pairs / t_all
    rate1 = pairs1 / t_one
    return {"value": rate, "unit": "pairs/s", "cores": threads, "kind": "port",
            **host_cpu(),
            "one_thread": {"value": rate1, "unit": "pairs/s", "sample_leaves": int(k1),
                           "sample_pairs": int(pairs1), "median_s": t_one},
            "tree_build_s": t_tree, "tree_build_threads": 1,
            "lists_kernel_median_s": t_all,
            "full_step_estimate_s": t_tree + total_pairs / rate,
            "sampled_fraction_of_pairs": pairs / total_pairs if total_pairs else 0.0,
            "sample": (f"oracle (CPU restatement of the reference path): lists + float kernel "
                       f"(nbody1 restatement, powf law) on {k} random listed leaves of the "
                       f"same tree ({pairs:.3e} pairs = {100 * pairs / total_pairs:.2f}% of the "
                       f"step's), median of 5 after a warm-up on {threads} OpenMP threads; "
                       f"one-thread figure on {k1} leaves; tree build of the whole load once, "
                       f"single-threaded ({t_tree:.1f} s)")}


def fp64_error(ctx, n):
    """The float forces of the last step's lists against the device FP64
    evaluation of the same lists (equal to the CPU FP64 oracle): relative
    RMS over the listed particles and the largest relative error where |F|
    exceeds 1e-3 of its RMS."""
    fx, fy, fz = ctx.forces()
    dx, dy, dz = ctx.forces_f64()
    nodes = ctx.nodes()
    leaves = ctx.get_lists()[0]
    listed = np.zeros(n + 1, np.int32)
    np.add.at(listed, nodes["offset"][leaves], 1)
    np.add.at(listed, nodes["offset"][leaves] + nodes["count"][leaves], -1)
    sel = np.cumsum(listed[:n]) > 0
    f = np.stack([fx, fy, fz], 1)[sel].astype(np.float64)
    d = np.stack([dx, dy, dz], 1)[sel]
    e = np.linalg.norm(f - d, axis=1)
    nrm = np.linalg.norm(d, axis=1)
    big = nrm > 1e-3 * np.sqrt(np.mean(nrm ** 2))
    return {"rel_rms": float(np.sqrt(np.sum(e ** 2) / np.sum(nrm ** 2))),
            "max_rel": float((e[big] / nrm[big]).max()) if big.any() else 0.0,
            "particles": int(sel.sum()), "reference": "device FP64 evaluation of the same lists"}


def build_roofline(ctx, n, build_ms):
    """Tree build against the HBM roofline with its algorithmic bytes."""
    nodes = ctx.nodes()
    internal = (nodes["cl"] != 0) | (nodes["cr"] != 0)
    split_particles = int(nodes["count"][internal].sum())
    nbytes = BUILD_BYTES_PER_LEVEL * split_particles + BUILD_BYTES_GATHER * n
    gbps = nbytes / (build_ms * 1e-3) / 1e9
    return {"tree_build": {"algorithmic_bytes": nbytes, "ms": build_ms, "GB/s": gbps,
                           "frac": gbps / PEAK_HBM_GBPS,
                           "bytes_rule": "48 B x particles per split level + 84 B x N gather"}}


def adjoint_leg(ctx, n, dev, reps, fw_stats):
    """hsr_adjoint (d(sum lambda.F)/dx, frozen topology) on the resident tree
    and lists of the last step, with a seeded Gaussian lambda.  The first call
    also builds the transposed lists (once per list build); the ratio compares
    the adjoint passes with the forward force phase on the same lists."""
    import torch
    g = torch.Generator().manual_seed(11)
    lam = [torch.randn(n, generator=g).to(dev) for _ in range(3)]
    out = [torch.empty(n, dtype=torch.float32, device=dev) for _ in range(3)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.adjoint(*lam, *out, device_ptrs=True)
    first_ms = 1e3 * (time.perf_counter() - t0)
    ms = []
    for _ in range(reps):
        ctx.adjoint(*lam, *out, device_ptrs=True)
        ms.append(ctx.stats().adjoint_ms)
    adj_ms = float(np.mean(ms))
    fw_ms = float(np.mean([s[3] for s in fw_stats]))
    fw_kernel_ms = float(np.mean([s[4] for s in fw_stats]))
    return {"adjoint_ms": adj_ms, "forward_ms": fw_ms, "forward_kernel_ms": fw_kernel_ms,
            "ratio": adj_ms / fw_ms, "first_call_ms_with_transpose": first_ms,
            "note": "adjoint_ms: target, source and pseudo passes (HIP events); forward_ms: "
                    "force phase of the same lists (fill, leaf kernel, chunk reduction)"}


def subcycle_leg(load, nd, nsub, dev):
    """Particles::subCycle on the device (hsr_subcycle) over the same load:
    nsub x (drift, resort, unit masses, tree, lists, kick, drift) with the
    particles resident for the whole cycle.  Seeded Gaussian velocities (sigma
    5 grid units per unit time) with a TimeStepper of pp 0.5, tau = tau2 =
    0.02, adot 1.3, fscal 1.1, alpha 1; grid ng = ceil(box + 2 overload) per
    side, edge = overload.  Drift and resort are HBM passes: drift moves 36 B
    per particle (reads x, y, z, vx, vy, vz; writes x, y, z), the resort at
    least 84 B (reads and writes the ten arrays, 42 B per particle)."""
    import math
    import torch
    from haccsr.particles import Particles, TimeStepperValues
    n = load.n
    ng = int(math.ceil(load.box + 2 * load.overload))
    g = torch.Generator().manual_seed(23)
    d = {"x": torch.from_numpy(load.x).to(dev), "y": torch.from_numpy(load.y).to(dev),
         "z": torch.from_numpy(load.z).to(dev)}
    for k in ("vx", "vy", "vz"):
        d[k] = (5.0 * torch.randn(n, generator=g)).to(dev)
    d["mass"] = torch.ones(n, device=dev)
    d["phi"] = torch.zeros(n, device=dev)
    d["id"] = torch.arange(n, dtype=torch.int64, device=dev)
    d["mask"] = torch.zeros(n, dtype=torch.int16, device=dev)
    ts = TimeStepperValues(pp=0.5, tau=0.02, tau2=0.02, adot=1.3, fscal=1.1)
    P = Particles([ng] * 3, alpha=1.0, gpscal=1.0, edge=load.overload, nsub=nsub,
                  rcb_tree_ppn=nd)
    P.ctx.set_tuning(use_graph=1)
    P.set_particles(d["x"], d["y"], d["z"], d["vx"], d["vy"], d["vz"], d["mass"], d["phi"],
                    d["id"], d["mask"])
    P.nsub = 1
    P.subCycle(ts)          # allocations and plans
    P.nsub = nsub
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    P.subCycle(ts)
    torch.cuda.synchronize()
    wall = 1e3 * (time.perf_counter() - t0)
    st = P.ctx.stats()
    P.close()
    drift_bytes, resort_bytes = 36 * n, 84 * n
    dr_ms, rs_ms = st.drift_ms / (2 * nsub), st.resort_ms / nsub
    return {"nsub": nsub, "ms": wall, "ms_per_substep": wall / nsub, "particles": n,
            "inside_last": P.Np_last, "ng": ng,
            "drift": {"ms": dr_ms, "algorithmic_bytes": drift_bytes,
                      "GB/s": drift_bytes / (dr_ms * 1e-3) / 1e9,
                      "frac": drift_bytes / (dr_ms * 1e-3) / 1e9 / PEAK_HBM_GBPS},
            "resort": {"ms": rs_ms, "algorithmic_bytes": resort_bytes,
                       "GB/s": resort_bytes / (rs_ms * 1e-3) / 1e9,
                       "frac": resort_bytes / (rs_ms * 1e-3) / 1e9 / PEAK_HBM_GBPS},
            "note": "one hsr_subcycle call, particles resident on the device; drift and "
                    "resort times are HIP-event means per pass"}


def pm_leg(load, dev, reps=3):
    """Particles::cic and Particles::inverse_cic on the same load (hsr_cic,
    hsr_inverse_cic; device arrays; grid ng = ceil(box + 2 overload) per
    side).  Algorithmic bytes: the deposit reads 12 B per particle and writes
    4 B per cell; the interpolation reads 12 B of position and 4 B of
    velocity and writes 4 B per particle, and reads the grid (4 B per cell)."""
    import math
    import torch
    from haccsr import Context
    n = load.n
    ng = int(math.ceil(load.box + 2 * load.overload))
    ncell = ng ** 3 + 1
    d = [torch.from_numpy(a).to(dev) for a in (load.x, load.y, load.z)]
    rho = torch.empty(ncell, dtype=torch.float32, device=dev)
    v = torch.zeros(n, dtype=torch.float32, device=dev)
    ctx = Context()
    ctx.cic(*d, [ng] * 3, 1.0, rho, device_ptrs=True)      # allocations
    ctx.inverse_cic(*d, [ng] * 3, rho, 0.01, 1.0, v, device_ptrs=True)
    cic_ms, icic_ms = [], []
    for _ in range(reps):
        ctx.cic(*d, [ng] * 3, 1.0, rho, device_ptrs=True)
        cic_ms.append(ctx.stats().cic_ms)
        ctx.inverse_cic(*d, [ng] * 3, rho, 0.01, 1.0, v, device_ptrs=True)
        icic_ms.append(ctx.stats().inverse_cic_ms)
    ctx.close()
    c_ms, i_ms = float(np.mean(cic_ms)), float(np.mean(icic_ms))
    cb = 12 * n + 4 * ncell
    ib = 20 * n + 4 * ncell
    return {"ng": ng, "particles": n,
            "cic": {"ms": c_ms, "algorithmic_bytes": cb, "GB/s": cb / (c_ms * 1e-3) / 1e9,
                    "frac": cb / (c_ms * 1e-3) / 1e9 / PEAK_HBM_GBPS},
            "inverse_cic": {"ms": i_ms, "algorithmic_bytes": ib,
                            "GB/s": ib / (i_ms * 1e-3) / 1e9,
                            "frac": ib / (i_ms * 1e-3) / 1e9 / PEAK_HBM_GBPS},
            "note": "deposit bit-exact with the reference's sequential float order: particles "
                    "sorted by base cell, each cell's eight neighbour lists merged by particle "
                    "index and its terms added in that order"}


def refresh_leg(load, dev, reps=3):
    """MC3Extras' particle refresh on one rank (hsr_refresh: alive particles
    kept, the 26 neighbours' overload copies appended, unit masses, resort;
    device arrays) on the same particles, whose old ghosts it drops and
    regenerates.  Domain: rL = ng = L, oL = the overload width (rounded up to
    whole cells by Domain::initialize).  Algorithmic bytes: the exchange
    reads 12 B per particle to classify and 38 B to pack, and writes 42 B per
    refreshed particle; the resort reads 12 B per refreshed particle for its
    keys and moves 84 B in its gather (the radix passes over the 8 B
    (cell, index) pairs are not counted)."""
    import torch
    from haccsr import Context, _lib
    ng = int(round(load.box))
    dom = _lib.domain(load.box, load.overload, ng)
    shift = float(dom.ng_overload) - float(load.overload)
    n = load.n
    cap = n + n // 4
    rng = np.random.default_rng(7)
    host = {"x": (load.x + shift).astype(np.float32), "y": (load.y + shift).astype(np.float32),
            "z": (load.z + shift).astype(np.float32)}
    for k in ("vx", "vy", "vz"):
        host[k] = rng.normal(size=n).astype(np.float32)
    host["mass"] = np.ones(n, np.float32)
    host["phi"] = np.zeros(n, np.float32)
    host["id"] = np.arange(n, dtype=np.int64)
    host["mask"] = np.zeros(n, np.uint16)
    names = ("x", "y", "z", "vx", "vy", "vz", "mass", "phi", "id", "mask")
    src = {k: torch.from_numpy(host[k]).to(dev) for k in names}
    buf = {k: torch.empty(cap, dtype=src[k].dtype, device=dev) for k in names}
    ctx = Context()
    ms, res = [], None
    for r in range(reps + 1):       # the first call allocates
        for k in names:
            buf[k][:n].copy_(src[k])
        torch.cuda.synchronize()
        res = ctx.refresh(dom, 1.0, n, *(buf[k] for k in names), device_ptrs=True)
        if r:
            ms.append(ctx.stats().refresh_ms)
    st = ctx.stats()
    ctx.close()
    m = res[0]
    t = float(np.mean(ms))
    b = 50 * n + 138 * m
    return {"ms": t, "particles_in": n, "alive": int(st.refresh_alive),
            "copies": int(st.refresh_dead), "inside": res[1], "ng_local_total": dom.ng_local_total,
            "algorithmic_bytes": b, "GB/s": b / (t * 1e-3) / 1e9,
            "frac": b / (t * 1e-3) / 1e9 / PEAK_HBM_GBPS,
            "note": "bit-exact with the oracle's restatement of copyAliveIntoVectors, "
                    "ParticleExchange (every neighbour this rank), copyAllFromVectors and "
                    "resortParticles"}


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: the short-range path is "
                          "per-rank with no communication inside it; see DESIGN.md for "
                          "how it shards", "n_gpus": a.gpus}), flush=True)
        sys.exit(2)
    import torch
    from haccsr import Context, autotune, ics

    if a.calibrate_copy:
        src = torch.ones(a.calibrate_copy // 4, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copied {a.calibrate_copy} bytes")
        return

    t0 = time.perf_counter()
    if a.config == "E":
        loads = ics.ensemble_loads(log=log)
        ens = {"offsets": np.cumsum([0] + [b.n for b in loads]).astype(np.int64),
               "lo": np.stack([b.tree_box[0] for b in loads]),
               "hi": np.stack([b.tree_box[1] for b in loads]),
               "flo": np.stack([b.force_box[0] for b in loads]),
               "fhi": np.stack([b.force_box[1] for b in loads]),
               "alive": sum(b.n_alive for b in loads), "nboxes": len(loads),
               "sides": [int(round(b.box)) for b in loads]}
        px = np.concatenate([b.x for b in loads])
        py = np.concatenate([b.y for b in loads])
        pz = np.concatenate([b.z for b in loads])
        n = px.shape[0]
        load = None
        log(f"load E: {ens['nboxes']} boxes, {n} particles ({ens['alive']} alive) "
            f"in {time.perf_counter() - t0:.1f} s")
    else:
        ens = None
        load = ics.config_load(a.config, n_side=a.n_side)
        log(f"load {a.config}: {load.n} particles ({load.n_alive} alive) in {time.perf_counter() - t0:.1f} s")
        n = load.n
        px, py, pz = load.x, load.y, load.z
        lo, hi = load.tree_box
        flo, fhi = load.force_box
    dev = torch.device("cuda:0")

    # the ten arrays of the unsorted load, and the working copy the step
    # restores with one device copy on the library's stream (the build
    # permutes the arrays in place)
    host = {"x": px, "y": py, "z": pz, "m": np.ones(n, np.float32),
            "vx": np.zeros(n, np.float32), "vy": np.zeros(n, np.float32),
            "vz": np.zeros(n, np.float32), "phi": np.zeros(n, np.float32),
            "id": np.arange(n, dtype=np.int64), "mask": np.zeros(n, np.int16)}
    layout, nbytes = {}, 0
    for k, v in host.items():
        layout[k] = (nbytes, v.dtype)
        nbytes += (v.nbytes + 255) // 256 * 256
    src_all = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    work_all = torch.empty(nbytes, dtype=torch.uint8, device=dev)

    def views(buf):
        out = {}
        for k, (o, dt) in layout.items():
            tdt = {np.dtype(np.float32): torch.float32, np.dtype(np.int64): torch.int64,
                   np.dtype(np.int16): torch.int16}[np.dtype(dt)]
            out[k] = buf[o:o + n * np.dtype(dt).itemsize].view(tdt)
        return out
    src, work = views(src_all), views(work_all)
    for k, v in host.items():
        src[k].copy_(torch.from_numpy(np.ascontiguousarray(v)))
    torch.cuda.synchronize()
    fcoeff = 1.0
    ctxs = {}
    shape = {"i_per_lane": a.i_per_lane, "unroll": a.unroll, "tile": a.tile}

    def step(nd):
        ctx = ctxs.get(nd)
        if ctx is None:
            ctx = ctxs[nd] = Context(nd=nd, tdpts=a.tdpts)
            ctx.torch_stream = torch.cuda.ExternalStream(ctx.stream())
        want = dict(use_graph=1, force_pass=a.force_pass, **shape)
        if getattr(ctx, "applied", None) != want:  # a change drops the captured graph
            ctx.set_tuning(**want)
            ctx.applied = want
        with torch.cuda.stream(ctx.torch_stream):
            work_all.copy_(src_all)
        if ens is None:
            ctx.build(work["x"], work["y"], work["z"], work["m"], work["vx"], work["vy"],
                      work["vz"], work["phi"], work["id"], work["mask"],
                      min_loc=lo, max_loc=hi, device_ptrs=True)
            ctx.lists(flo, fhi)
        else:
            # every box's tree, lists and leaf work in the same launches
            ctx.build_many(ens["offsets"], work["x"], work["y"], work["z"], work["m"],
                           ens["lo"], ens["hi"], work["vx"], work["vy"], work["vz"],
                           work["phi"], work["id"], work["mask"], device_ptrs=True)
            ctx.lists_many(ens["flo"], ens["fhi"])
        ctx.kick(fcoeff)
        return ctx.stats()

    nd = a.nd
    tuned = {}
    if a.tdpts == 1 and not (a.i_per_lane or a.unroll or a.tile):
        # the kernel shape on the lists of one leaf size, then the leaf size
        # with that shape (the two interact weakly: the shape sets the
        # leaf kernel's speed per pair, the leaf size the pairs and the rest)
        nd0 = nd if nd > 0 else 512
        step(nd0)
        best, times = autotune.tune_kernel_shape(ctxs[nd0], lambda: ctxs[nd0].kick(fcoeff), log=log)
        shape.update(best)
        ctxs[nd0].applied = None
        tuned["kernel_shape"] = dict(best, tuned_at_nd=nd0,
                                     ms={f"{r},{u},{t}": round(v, 3) for (r, u, t), v in times.items()})
        log(f"autotuned kernel shape {best}")
    if nd <= 0:
        nd, times = autotune.tune_leaf_size(step, log=log)
        tuned["leaf_size"] = {"nd": nd, "ms_per_step": {str(k): round(1e3 * v, 2)
                                                        for k, v in sorted(times.items())}}
        log(f"autotuned leaf size nd={nd}")
    for k in list(ctxs):
        if k != nd:
            ctxs.pop(k).close()
    for _ in range(a.warmup):
        step(nd)
    torch.cuda.synchronize()
    g0 = ctxs[nd].stats() if a.warmup > 0 else None
    stats = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        st = step(nd)
        stats.append((st.pairs, st.build_ms, st.lists_ms, st.force_ms, st.kernel_ms,
                      st.evaluated_pairs))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = 1e3 * wall / a.steps
    pairs = stats[-1][0]
    kern_ms = float(np.mean([s[4] for s in stats]))
    evaluated = stats[-1][5]
    # the roofline counts the pairs the kernel evaluated (entries beyond the
    # cutoff of a whole leaf are skipped; the metric counts list pairs)
    achieved = FLOPS_PER_PAIR * evaluated / (kern_ms * 1e-3) / 1e12
    st = ctxs[nd].stats()
    log(f"pairs/step {pairs:.4e}  ms/step {ms_per_step:.2f}  build {np.mean([s[1] for s in stats]):.2f}"
        f"  lists {np.mean([s[2] for s in stats]):.2f}  kick {np.mean([s[3] for s in stats]):.2f}"
        f"  leaf kernel {kern_ms:.2f} ms  nodes {st.nodes} listed leaves {st.listed_leaves}"
        f"  max depth {st.max_depth}")
    out_shape = {"i_per_lane": shape["i_per_lane"] or 2, "unroll": shape["unroll"] or 2,
                 "tile": shape["tile"] or 256}
    out = {
        "metric": "pair interactions/s & % FP32 VALU peak at 256^3; adjoint/forward time",
        "value": pairs / (ms_per_step * 1e-3),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded clustered generator; unit masses)",
        "config": {"workload": (workload_name(a.config, load, nd) if ens is None else
                                ensemble_name(ens, n, nd)), "particles": n,
                   "pairs_per_step": pairs, "list_entries": int(st.entries), "leaf_size": nd,
                   "kernel_shape": out_shape,
                   "autotune": tuned,
                   "breakdown_ms": {"build": float(np.mean([s[1] for s in stats])),
                                    "lists": float(np.mean([s[2] for s in stats])),
                                    "kick": float(np.mean([s[3] for s in stats])),
                                    "leaf_kernel": kern_ms},
                   # the force phase is one captured graph, replayed across
                   # the steps' list builds
                   "force_graph": {"captures_in_timed_steps":
                                   int(st.graph_captures - (g0.graph_captures if g0 else 0)),
                                   "replays_in_timed_steps":
                                   int(st.graph_replays - (g0.graph_replays if g0 else 0))}},
        "roofline": {"bound": "valu", "pipe": "FP32 VALU (no MFMA: the pair force is not a contraction)",
                     "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS,
                     "traffic": MEASURED_TRAFFIC.get((a.config, a.n_side or 256, nd, a.tdpts,
                                                      tuple(out_shape.values()), KERNEL_SRC)),
                     "kernel": "k_force (leaf kernel)", "flops_per_pair": FLOPS_PER_PAIR,
                     "evaluated_pairs": evaluated, "unmasked_pairs": int(st.unmasked_pairs),
                     "culled_fraction_of_list_pairs": 1.0 - evaluated / pairs if pairs else 0.0},
        "hbm_passes": build_roofline(ctxs[nd], n, float(np.mean([s[1] for s in stats]))),
    }
    if a.tdpts != 1:
        out["config"]["workload"] += "; quadrupole tree (TDPTS 12)"
    out["roofline"]["kernel_source_sha16"] = KERNEL_SRC
    out["fp64_error"] = fp64_error(ctxs[nd], n)
    log(f"float vs FP64 {out['fp64_error']}")
    if a.adjoint_reps > 0 and a.tdpts == 1:
        ctxs[nd].set_tuning(adjoint_mode=a.adjoint_mode)
        out["adjoint"] = adjoint_leg(ctxs[nd], n, dev, a.adjoint_reps, stats)
        log(f"adjoint {out['adjoint']}")
    if a.subcycle_nsub > 0 and ens is None:
        for k in list(ctxs):
            ctxs.pop(k).close()
        out["subcycle"] = subcycle_leg(load, nd, a.subcycle_nsub, dev)
        log(f"subcycle {out['subcycle']}")
    if a.pm_reps > 0 and ens is None:
        out["pm"] = pm_leg(load, dev, a.pm_reps)
        log(f"pm {out['pm']}")
    if a.refresh_reps > 0 and ens is None:
        out["refresh"] = refresh_leg(load, dev, a.refresh_reps)
        log(f"refresh {out['refresh']}")
    if not a.no_cpu and ens is None:
        log("cpu baseline leg ...")
        out["cpu_baseline"] = cpu_baseline(load, nd)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
