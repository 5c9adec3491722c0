"""hsr_resort_noalloc, the allocation-free resort: through the C ABI on
device arrays, bit for bit the result of hsr_resort with device pointers (all
ten arrays and m_Np_last), on two calls in a row with the second smaller, and
its limits refused before any launch."""
import ctypes

import numpy as np
import pytest

from haccsr import Context, HsrError
from haccsr._lib import HSR_ERR_ARG, HSR_HOST_PTRS
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ARR = ("x", "y", "z", "vx", "vy", "vz", "mass", "phi", "id", "mask")
NOALLOC_MAX = 1 << 22  # HSR_RESORT_NOALLOC_MAX in include/hacc_sr.h


def say(msg):
    print(msg, flush=True)


def load(n, ng, seed):
    """Uniform particles with a shell outside the grid, plus clumps (long
    cell runs), in random order."""
    rng = np.random.default_rng(seed)
    g = np.asarray(ng, np.float64)
    nu = n // 2
    p = rng.uniform(0, 1, (nu, 3)) * (g + 3.0) - 1.5
    centres = rng.uniform(0.2, 0.8, (9, 3)) * g
    c = centres[rng.integers(0, 9, n - nu)] + rng.normal(0, 0.7, (n - nu, 3))
    pos = np.concatenate([p, c]).astype(np.float32)[rng.permutation(n)]
    st = {"x": pos[:, 0].copy(), "y": pos[:, 1].copy(), "z": pos[:, 2].copy()}
    for a in ("vx", "vy", "vz", "phi"):
        st[a] = rng.normal(0, 3, n).astype(np.float32)
    st["mass"] = rng.uniform(0.5, 2.0, n).astype(np.float32)
    st["id"] = rng.permutation(n).astype(np.int64) * 5 + 1
    st["mask"] = rng.integers(0, 65535, n).astype(np.uint16)
    return st


def to_dev(v):
    import torch
    a = v.copy()
    return torch.from_numpy(a.view(np.int16) if a.dtype == np.uint16 else a).cuda()


def to_host(t, like):
    a = t.cpu().numpy()
    return a.view(np.uint16) if like.dtype == np.uint16 else a


def bits(a):
    a = np.asarray(a)
    return a.view(np.uint32) if a.dtype == np.float32 else a


def run(ctx, fn, st, ng, present):
    dev = {a: (to_dev(st[a]) if a in present else None) for a in ARR}
    args = [dev[a] for a in ARR]
    inside = fn(ng, *args[:3], vx=args[3], vy=args[4], vz=args[5], mass=args[6], phi=args[7],
                ids=args[8], mask=args[9], **({"device_ptrs": True} if fn == ctx.resort else {}))
    return inside, {a: to_host(dev[a], st[a]) for a in present}


def test_noalloc_resort_matches_hsr_resort_twice():
    import torch
    assert torch.cuda.is_available()
    ctx = Context()
    cases = [
        # (n, ng, seed, arrays present): the second call is smaller, on another
        # grid and with fewer arrays, so stale scratch from the first would show
        (1_500_000, [96, 96, 96], 21, ARR),
        (300_007, [40, 56, 72], 22, ("x", "y", "z", "vx", "mass", "id")),
    ]
    for n, ng, seed, present in cases:
        say(f"case noalloc resort n={n} ng={ng} arrays={len(present)}")
        st = load(n, ng, seed)
        want_in, want = run(ctx, ctx.resort, st, ng, present)
        got_in, got = run(ctx, ctx.resort_noalloc, st, ng, present)
        assert got_in == want_in
        for a in present:
            assert np.array_equal(bits(got[a]), bits(want[a])), a
        # and the permutation is the reference's cell sort (the oracle's
        # perm[i] is particle i's destination)
        perm, o_in = O.cell_sort_perm(st["x"], st["y"], st["z"], ng)
        want_id = np.empty_like(st["id"])
        want_id[perm] = st["id"]
        assert o_in == got_in and np.array_equal(got["id"], want_id)
    # a repeat of the second case after the first left its scratch behind
    n, ng, seed, present = cases[1]
    st = load(n, ng, seed)
    want_in, want = run(ctx, ctx.resort, st, ng, present)
    got_in, got = run(ctx, ctx.resort_noalloc, st, ng, present)
    assert got_in == want_in and all(np.array_equal(bits(got[a]), bits(want[a])) for a in present)
    ctx.close()


def test_noalloc_resort_limits():
    import torch
    ctx = Context()
    say("case noalloc resort limits")
    # host arrays are refused (they would need staging that hsr_resort allocates)
    st = load(1000, [8, 8, 8], 3)
    ng = np.asarray([8, 8, 8], np.int32)
    inside = ctypes.c_int64(-1)
    status = ctx.L.hsr_resort_noalloc(ctx.h, 1000, ng.ctypes.data, st["x"].ctypes.data,
                                      st["y"].ctypes.data, st["z"].ctypes.data,
                                      *([None] * 7), ctypes.byref(inside), HSR_HOST_PTRS)
    assert status == HSR_ERR_ARG
    assert "device pointers only" in ctx.L.hsr_last_error().decode()
    # one particle above the limit is refused before any launch
    big = [torch.zeros(NOALLOC_MAX + 1, dtype=torch.float32, device="cuda") for _ in range(3)]
    with pytest.raises(HsrError, match="limit"):
        ctx.resort_noalloc([16, 16, 16], *big)
    # exactly the limit is accepted and agrees with hsr_resort
    say("case noalloc resort n=max")
    st = load(NOALLOC_MAX, [128, 128, 128], 5)
    want_in, want = run(ctx, ctx.resort, st, [128, 128, 128], ("x", "y", "z", "id"))
    got_in, got = run(ctx, ctx.resort_noalloc, st, [128, 128, 128], ("x", "y", "z", "id"))
    assert got_in == want_in
    for a in ("x", "y", "z", "id"):
        assert np.array_equal(bits(got[a]), bits(want[a])), a
    ctx.close()

