"""Pins of the oracle's one-rank refresh (hsro_domain_init, hsro_refresh:
Domain.cxx:34-84, MC3Extras.cxx:660-700, ParticleExchange.cxx:103-728,
Particles.cxx:298-366, 762-830, 945-1030).

The reference's ParticleExchange needs MPI (Partition.h) and cannot be built
here, and its tests hold no vectors; the oracle is pinned by what the
mathematics fixes instead:

* the Domain values, as known answers worked out by hand;
* the copies: with one rank they must be exactly the periodic images
  p + k rL, k in {-1, 0, 1}^3 minus 0, of the alive particles p whose every
  shifted dimension lies in the closed overload shell (k_d = +1 needs
  p_d <= oL, k_d = -1 needs p_d >= rL - oL), enumerated here by brute force
  over k; their order is neighbour by neighbour in the enum order of
  BasicDefinition.h:224-264, particle order within one, where the neighbour
  of shift k is named by the faces it crosses (X0 is the low x face, whose
  copies move by +rL);
* alive particles keep their input order, every mass is 1, and the unit
  conversions are identities where they are exact (rL = ng).
"""
import itertools

import numpy as np
import pytest

from oracle import oracle as O

NAMES = ["X0", "X1", "Y0", "Y1", "Z0", "Z1", "X0_Y0", "X1_Y1", "X0_Y1", "X1_Y0",
         "Y0_Z0", "Y1_Z1", "Y0_Z1", "Y1_Z0", "Z0_X0", "Z1_X1", "Z0_X1", "Z1_X0",
         "X0_Y0_Z0", "X1_Y1_Z1", "X0_Y0_Z1", "X1_Y1_Z0", "X0_Y1_Z0", "X1_Y0_Z1",
         "X0_Y1_Z1", "X1_Y0_Z0"]


def shift_of(name):
    """The shift (in units of rL) of the copies sent across the named faces:
    a low face ('0') sends to the rank below, whose copies sit beyond this
    box's high side (+1), a high face ('1') the other way (-1)."""
    k = [0, 0, 0]
    for part in name.split("_"):
        k["XYZ".index(part[0])] = 1 if part[1] == "0" else -1
    return tuple(k)


def test_domain_known_answers():
    d = O.domain(256.0, 3.2, 256)
    assert (d.phys2grid_pos, d.grid2phys_pos, d.ng_overload, d.oL, d.ng_local_total) == \
        (1.0, 1.0, 4, 4.0, 265)
    assert d.grid2phys_vel == 100.0 and d.phys2grid_vel == np.float32(0.01)
    # 512 / 1000 = 0.512 (float), 10 * 0.512 -> 6 cells of 1.953125
    d = O.domain(1000.0, 10.0, 512)
    assert d.phys2grid_pos == np.float32(0.512) and d.grid2phys_pos == 1.953125
    assert (d.ng_overload, d.oL, d.ng_local_total) == (6, 11.71875, 525)
    assert d.grid2phys_vel == 195.3125 and d.phys2grid_vel == np.float32(0.00512)
    d = O.domain(64.0, 0.0, 64)
    assert (d.ng_overload, d.oL, d.ng_local_total) == (0, 0.0, 65)


def state_of(phys, ng_overload, seed=0):
    """Local grid coordinates (rL = ng: local = physical + ng_overload) and
    the other nine arrays."""
    rng = np.random.default_rng(seed)
    n = phys.shape[0]
    loc = (phys + np.float32(ng_overload)).astype(np.float32)
    st = dict(x=loc[:, 0].copy(), y=loc[:, 1].copy(), z=loc[:, 2].copy(),
              vx=rng.normal(size=n).astype(np.float32),
              vy=rng.normal(size=n).astype(np.float32),
              vz=rng.normal(size=n).astype(np.float32),
              mass=rng.uniform(0.5, 2, n).astype(np.float32),
              phi=rng.normal(size=n).astype(np.float32),
              id=np.arange(n, dtype=np.int64) * 7 + 3,
              mask=rng.integers(0, 65535, n).astype(np.uint16))
    return st


def padded(st, cap):
    out = {}
    for k, a in st.items():
        b = np.zeros(cap, a.dtype)
        b[:a.shape[0]] = a
        out[k] = b
    return out


def boundary_points(rL, oL):
    """Every combination of the values where the alive and shell tests
    switch, plus random points inside and outside the box."""
    vals = np.array([-0.5, 0.0, 0.25, oL, oL + 0.125, rL / 2, rL - oL - 0.125, rL - oL,
                     rL - 0.25, rL, rL + 0.5], np.float32)
    grid = np.array(list(itertools.product(vals, vals, vals)), np.float32)
    rng = np.random.default_rng(5)
    rand = rng.uniform(-oL, rL + oL, (3000, 3)).astype(np.float32)
    return np.concatenate([grid, rand])


@pytest.mark.parametrize("rL,oL", [(64, 4.0), (32, 3.0)])
def test_refresh_copies_are_the_periodic_images(rL, oL):
    d = O.domain(rL, oL, rL)
    assert d.oL == oL and d.phys2grid_pos == 1.0
    phys = boundary_points(rL, oL)
    st = state_of(phys, d.ng_overload)
    n = phys.shape[0]
    cap = 8 * n
    # unsorted output: call the C restatement directly
    s2 = padded(st, cap)
    out, count, inside, alive = O.refresh(s2, n, rL, oL, rL, 1.0, sort=False)
    # physical coordinates as the reference's float conversion sees them
    # (exact: a float x >= 4 minus 4 is representable)
    og = np.float32(d.ng_overload)
    phys = np.stack([st[k] for k in "xyz"], 1) - og
    # alive particles, in input order
    keep = np.all((phys >= 0) & (phys < rL), axis=1)
    assert alive == int(keep.sum())
    ids = st["id"][keep]
    assert np.array_equal(out["id"][:alive], ids)
    for k, c in zip("xyz", range(3)):
        assert np.array_equal(out[k][:alive], st[k][keep])
    # brute-force images, ordered by the neighbour of each shift
    order = {shift_of(nm): i for i, nm in enumerate(NAMES)}
    assert sorted(order) == sorted(k for k in itertools.product((-1, 0, 1), repeat=3)
                                   if k != (0, 0, 0))
    P = phys[keep]
    want = []
    for k, nb in sorted(order.items(), key=lambda t: t[1]):
        ok = np.ones(P.shape[0], bool)
        for dim in range(3):
            if k[dim] == 1:
                ok &= P[:, dim] <= np.float32(oL)
            elif k[dim] == -1:
                ok &= P[:, dim] >= np.float32(rL) - np.float32(oL)
        for i in np.nonzero(ok)[0]:
            want.append((ids[i],) + tuple((P[i, dim] + np.float32(k[dim] * rL)) + og
                                          for dim in range(3)))
    assert count == alive + len(want)
    got = [(out["id"][j], out["x"][j], out["y"][j], out["z"][j]) for j in range(alive, count)]
    assert got == want
    # unit masses; velocities through 100/a^2 and a^2/100 (anow = 1)
    assert np.all(out["mass"][:count] == 1.0)
    src = {i: j for j, i in enumerate(st["id"])}
    rows = np.array([src[i] for i in out["id"][:count]])
    for k in ("vx", "vy", "vz"):
        assert np.allclose(out[k][:count], st[k][rows], rtol=3e-7, atol=0)
    assert np.array_equal(out["phi"][:count], st["phi"][rows])
    assert np.array_equal(out["mask"][:count], st["mask"][rows])
    # the sorted variant is the cell sort of the same records
    s3, count3, inside3, alive3 = O.refresh(padded(st, cap), n, rL, oL, rL, 1.0)
    assert (count3, alive3) == (count, alive)
    perm, ins = O.cell_sort_perm(out["x"][:count], out["y"][:count], out["z"][:count],
                                 [d.ng_local_total] * 3)
    assert ins == inside3
    assert np.array_equal(s3["id"][perm], out["id"][:count])


def test_refresh_capacity_and_empty():
    d = O.domain(16, 2.0, 16)
    phys = np.array([[1.0, 1.0, 1.0]], np.float32)   # a corner: 7 copies
    st = state_of(phys, d.ng_overload)
    with pytest.raises(ValueError):
        O.refresh(padded(st, 7), 1, 16, 2.0, 16, 1.0)
    out, count, inside, alive = O.refresh(padded(st, 8), 1, 16, 2.0, 16, 1.0)
    assert (count, alive) == (8, 1)
    out, count, inside, alive = O.refresh(padded(st, 8), 0, 16, 2.0, 16, 1.0)
    assert (count, alive, inside) == (0, 0, 0)


def test_refresh_units_with_scale_factor():
    """rL != ng and anow != 1: positions of alive particles go through
    (x - oLg) * rL/ng * ng/rL + oLg and come back within an ulp or two; a
    copy lands rL away, i.e. ng grid units."""
    rL, oL, ng, a = 100.0, 8.0, 64, 0.37
    d = O.domain(rL, oL, ng)
    rng = np.random.default_rng(2)
    loc = rng.uniform(0, ng + 2 * d.ng_overload, (20000, 3)).astype(np.float32)
    st = state_of(loc - np.float32(d.ng_overload), d.ng_overload, seed=3)
    n = loc.shape[0]
    out, count, inside, alive = O.refresh(padded(st, 4 * n), n, rL, oL, ng, a, sort=False)
    src = {i: j for j, i in enumerate(st["id"])}
    first = {}
    for j in range(count):
        first.setdefault(out["id"][j], j)
    for j in range(count):
        i = src[out["id"][j]]
        for k in "xyz":
            dx = float(out[k][j]) - float(st[k][i])
            assert min(abs(dx), abs(abs(dx) - ng)) < 1e-4 * ng, (j, k, dx)
        for k in ("vx", "vy", "vz"):
            assert abs(out[k][j] - st[k][i]) <= 1e-6 * abs(st[k][i]) + 1e-30
