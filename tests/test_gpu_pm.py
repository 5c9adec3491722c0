"""The PM particle-grid steps on the device (hsr_cic, hsr_inverse_cic):
bit-exact with the oracle's restatement of Particles::cic and
Particles::inverse_cic, including the reference's sequential order of float
additions into each cell and the slot for corners outside the grid."""
import numpy as np
import pytest

from haccsr import Context, HsrError, ics
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def positions(n, ng, seed, clump=0.3):
    """Uniform over the grid and a margin outside it, plus tight clumps so
    that hundreds of particles share cells."""
    rng = np.random.default_rng(seed)
    g = np.asarray(ng, np.float64)
    nu = int(n * (1 - clump))
    p = rng.uniform(0, 1, (nu, 3)) * (g + 3) - 1.5
    cen = rng.uniform(0.2, 0.8, (5, 3)) * g
    q = cen[rng.integers(0, 5, n - nu)] + rng.normal(0, 0.05, (n - nu, 3))
    pos = np.concatenate([p, q]).astype(np.float32)[rng.permutation(n)]
    return pos[:, 0].copy(), pos[:, 1].copy(), pos[:, 2].copy()


@pytest.mark.parametrize("n,ng,seed", [(0, (4, 4, 4), 1), (1, (4, 5, 6), 2),
                                       (50000, (16, 12, 20), 3), (2000000, (64, 64, 64), 4)])
def test_cic_deposit_bitexact(n, ng, seed):
    say(f"case cic n={n} ng={ng}")
    x, y, z = positions(max(n, 1), ng, seed)
    x, y, z = x[:n], y[:n], z[:n]
    want = O.cic(x, y, z, ng, gpscal=1.25)
    ctx = Context()
    got = ctx.cic(x, y, z, ng, gpscal=1.25)
    say(f"  {ctx.stats().cic_ms:.3f} ms")
    ctx.close()
    assert got.shape == want.shape
    bad = np.nonzero(bits(got) != bits(want))[0]
    assert bad.size == 0, f"cells differ: {bad[:5]} ({got[bad[:3]]} vs {want[bad[:3]]})"


@pytest.mark.parametrize("order", ["cell", "mixed"])
def test_cic_deposit_orders(order):
    """The deposit merges each cell's eight neighbour lists by particle
    index: particles in cell order (as the resort leaves them) make the lists
    disjoint index ranges, concatenated in order; a cell order with every
    8th particle swapped to a random place leaves some dense cells' lists
    interleaved, which are merged one contribution at a time.  Both
    bit-exact, dense clumps (long cells) and the outside slot included."""
    say(f"case cic input order {order}")
    ng = (40, 36, 44)
    x, y, z = positions(600000, ng, 11, clump=0.5)
    ix, iy, iz = (np.floor(a).astype(np.int64) for a in (x, y, z))
    inb = (ix >= 0) & (ix < ng[0]) & (iy >= 0) & (iy < ng[1]) & (iz >= 0) & (iz < ng[2])
    key = np.where(inb, (ix * ng[1] + iy) * ng[2] + iz, ng[0] * ng[1] * ng[2])
    perm = np.argsort(key, kind="stable")
    if order == "mixed":
        rng = np.random.default_rng(12)
        sel = np.arange(0, perm.shape[0], 8)
        perm[sel] = perm[sel][rng.permutation(sel.shape[0])]
    x, y, z = x[perm].copy(), y[perm].copy(), z[perm].copy()
    want = O.cic(x, y, z, ng, gpscal=1.25)
    ctx = Context()
    got = ctx.cic(x, y, z, ng, gpscal=1.25)
    ctx.close()
    bad = np.nonzero(bits(got) != bits(want))[0]
    assert bad.size == 0, f"cells differ: {bad[:5]} ({got[bad[:3]]} vs {want[bad[:3]]})"


def test_cic_device_tensors_and_reuse():
    """Device arrays in place; a second call on the same grid starts from
    zero again."""
    import torch
    say("case cic device tensors")
    ng = (32, 32, 32)
    x, y, z = positions(300000, ng, 7)
    want = O.cic(x, y, z, ng)
    d = [torch.from_numpy(a).cuda() for a in (x, y, z)]
    rho = torch.full((32 ** 3 + 1,), 7.0, device="cuda")
    ctx = Context()
    for _ in range(2):
        ctx.cic(*d, ng, 1.0, rho, device_ptrs=True)
        torch.cuda.synchronize()
        assert np.array_equal(bits(rho.cpu().numpy()), bits(want))
    ctx.close()


@pytest.mark.parametrize("n,ng,seed", [(1, (4, 5, 6), 5), (100000, (20, 24, 16), 6)])
def test_inverse_cic_bitexact(n, ng, seed):
    say(f"case inverse cic n={n} ng={ng}")
    x, y, z = positions(n, ng, seed)
    rng = np.random.default_rng(seed)
    grid = rng.normal(size=int(np.prod(ng)) + 1).astype(np.float32)
    grid[-1] = 55.0    # read as 0
    v0 = rng.normal(size=n).astype(np.float32)
    want = O.inverse_cic(x, y, z, ng, grid, 0.013, 1.7, v0)
    v = v0.copy()
    ctx = Context()
    ctx.inverse_cic(x, y, z, ng, grid, 0.013, 1.7, v)
    ctx.close()
    assert np.array_equal(bits(v), bits(want))


def test_inverse_cic_device_and_potential():
    """inverse_cic_potential's pattern (phi zeroed, tau = fscal = 1) on
    device tensors."""
    import torch
    say("case inverse cic device, potential")
    ng = (24, 24, 24)
    x, y, z = positions(200000, ng, 8)
    grid = O.cic(x, y, z, ng)          # any field; the deposit of the same particles
    want = O.inverse_cic(x, y, z, ng, grid, 1.0, 1.0, np.zeros_like(x))
    d = [torch.from_numpy(a).cuda() for a in (x, y, z)]
    g = torch.from_numpy(grid).cuda()
    phi = torch.zeros_like(d[0])
    ctx = Context()
    ctx.inverse_cic(*d, ng, g, 1.0, 1.0, phi, device_ptrs=True)
    torch.cuda.synchronize()
    assert np.array_equal(bits(phi.cpu().numpy()), bits(want))
    ctx.close()


def test_cic_rejects_bad_grid():
    ctx = Context()
    x = np.zeros(3, np.float32)
    with pytest.raises(HsrError):
        ctx.cic(x, x, x, (0, 2, 2))
    ctx.close()
