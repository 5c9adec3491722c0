"""Leaf-size and kernel-shape autotuner.

Two knobs decide the speed of a short-range step on a given load:

* the leaf size nd, which shapes the tree and therefore the interaction
  lists: larger leaves mean fewer, longer lists (more pairs, less tree and
  list work, longer LDS tiles).  A tuned nd is applied identically to the
  device run and to any comparison against the oracle;
* the leaf kernel's shape: i-particles per lane, j entries per unrolled step
  and entries per LDS tile, chosen among the instantiations built into the
  library (hsr_tuning i_per_lane / unroll / tile).  The shape changes only
  the order of the float sums.

tune_kernel_shape times the force phase of the lists resident in a context
under every built shape; tune_leaf_size times whole steps (build + lists +
kick) over leaf sizes until the fastest is bracketed by slower neighbours on
both sides, then tries the midpoints next to it.
"""
from __future__ import annotations

import time

# (i_per_lane, unroll, tile) built into libhaccsr.so (include/hacc_sr.h,
# hsr_tuning); the first is the library default
SHAPES = ((2, 2, 256), (2, 1, 256), (2, 4, 256), (1, 2, 256), (1, 4, 256), (4, 1, 256),
          (2, 2, 128), (2, 2, 512))

DEFAULT_CANDIDATES = (128, 192, 256, 384, 512, 768, 1024)
ND_MIN, ND_MAX = 16, 8192


def tune_kernel_shape(ctx, run_force, shapes=SHAPES, repeats=2, log=None):
    """Time run_force() (one force phase on ctx's current lists) under each
    shape by the leaf kernel's device time (hsr_stats.kernel_ms); leave ctx
    tuned to the fastest.  Returns ({"i_per_lane", "unroll", "tile"},
    {shape: ms})."""
    times = {}
    for r, u, t in shapes:
        ctx.set_tuning(i_per_lane=r, unroll=u, tile=t)
        run_force()  # warm-up (graph capture, first launch of the instantiation)
        best = float("inf")
        for _ in range(repeats):
            run_force()
            best = min(best, ctx.stats().kernel_ms)
        times[(r, u, t)] = best
        if log:
            log(f"autotune shape (i_per_lane {r}, unroll {u}, tile {t}): {best:.2f} ms")
    r, u, t = min(times, key=times.get)
    ctx.set_tuning(i_per_lane=r, unroll=u, tile=t)
    return {"i_per_lane": r, "unroll": u, "tile": t}, times


TIE = 0.005  # leaf sizes within half a percent of the fastest count as tied


def tune_leaf_size(run_step, candidates=DEFAULT_CANDIDATES, repeats=2, log=None, refine=True):
    """run_step(nd) performs one synchronized step with leaf size nd.
    Returns (best_nd, {nd: seconds}).  Candidates are extended by factors of
    1.5 past an edge until the fastest has a slower neighbour on each side
    (or a bound is reached); with refine, the midpoints between the fastest
    and its neighbours are timed too.  The step time is flat near its
    minimum, so the smallest leaf size within TIE of the fastest is chosen:
    run-to-run noise then does not move the choice."""
    times = {}

    def timeit(nd):
        if nd in times:
            return times[nd]
        run_step(nd)  # warm-up (allocations, first-touch)
        best = float("inf")
        for _ in range(repeats):
            t0 = time.perf_counter()
            run_step(nd)
            best = min(best, time.perf_counter() - t0)
        times[nd] = best
        if log:
            log(f"autotune nd={nd}: {best * 1e3:.2f} ms/step")
        return best

    for nd in candidates:
        timeit(nd)
    while True:
        order = sorted(times)
        best = min(times, key=times.get)
        k = order.index(best)
        if k == len(order) - 1 and best < ND_MAX:
            timeit(min(ND_MAX, int(best * 1.5)))
        elif k == 0 and best > ND_MIN:
            timeit(max(ND_MIN, int(best / 1.5)))
        else:
            break
    if refine:
        order = sorted(times)
        best = min(times, key=times.get)
        k = order.index(best)
        for j in (k - 1, k + 1):
            if 0 <= j < len(order):
                mid = (order[j] + best) // 2
                if mid not in times and mid != order[j]:
                    timeit(mid)
    tmin = min(times.values())
    best_nd = min(nd for nd, t in times.items() if t <= tmin * (1 + TIE))
    return best_nd, times
