"""Leaf-size and kernel-shape autotuner.

The leaf size nd shapes the tree (and therefore the interaction lists), so a
tuned nd is applied identically to the device run and to any comparison
against the oracle.  The tuner times the whole step (build + lists + kick)
for each candidate on the actual load and keeps the fastest.
"""
from __future__ import annotations

import time

DEFAULT_CANDIDATES = (64, 96, 128, 160, 192, 256, 320, 384, 512)


def tune_leaf_size(run_step, candidates=DEFAULT_CANDIDATES, repeats=2, log=None):
    """run_step(nd) performs one synchronized step with leaf size nd.
    Returns (best_nd, {nd: seconds})."""
    times = {}
    for nd in candidates:
        run_step(nd)  # warm-up (allocations, first-touch)
        best = float("inf")
        for _ in range(repeats):
            t0 = time.perf_counter()
            run_step(nd)
            best = min(best, time.perf_counter() - t0)
        times[nd] = best
        if log:
            log(f"autotune nd={nd}: {best * 1e3:.2f} ms/step")
    best_nd = min(times, key=times.get)
    return best_nd, times
