"""Synthetic particle loads for the short-range force path.

The reference obtains its particles from the Zel'dovich initializer and the
ghost ("overload") exchange (src/initializer, ParticleExchange.cxx:141-229);
neither is on the short-range path.  These generators make seeded inputs of
the same shape for tests and the benchmark:

* ``plummer_mix``   70 % uniform + 30 % in Plummer clumps (the 64^3 parity load)
* ``clustered``     Gaussian random field P(k) ~ k^-2, first-order Zel'dovich
                    displacement scaled to a target RMS, plus NFW halos
* ``add_ghosts``    periodic images inside a shell of width ``o`` around the
                    box, in the 26-neighbour order of BasicDefinition.h:224-264,
                    then a shift by +o so the tree box is [0, L + 2o)
* ``cell_sort``     the stable counting sort by unit cell that
                    Particles::resortParticles applies before every tree
                    build (Particles.cxx:402-488), which fixes the tree's
                    input order.

All positions are float32 and masses are 1 (Particles.cxx:1256-1257).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

# Neighbour directions in the enum order of BasicDefinition.h:224-264.
# X0 is the low-x face, X1 the high-x face, etc.
_NEIGHBORS = [
    (-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1),
    (-1, -1, 0), (1, 1, 0), (-1, 1, 0), (1, -1, 0),
    (0, -1, -1), (0, 1, 1), (0, -1, 1), (0, 1, -1),
    (-1, 0, -1), (1, 0, 1), (1, 0, -1), (-1, 0, 1),
    (-1, -1, -1), (1, 1, 1), (-1, -1, 1), (1, 1, -1),
    (-1, 1, -1), (1, -1, 1), (-1, 1, 1), (1, -1, -1),
]


@dataclasses.dataclass
class Load:
    """One rank's particles laid out as the tree constructor receives them."""
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    n_alive: int          # the first n_alive entries are the box's own particles
    box: float            # periodic box side L
    overload: float       # ghost shell width o
    alive_index: np.ndarray  # for each entry, index of the alive particle it images

    @property
    def n(self) -> int:
        return int(self.x.shape[0])

    @property
    def tree_box(self):
        hi = np.float32(self.box + 2 * self.overload)
        return (np.zeros(3, np.float32), np.full(3, hi, np.float32))

    @property
    def force_box(self):
        lo = np.float32(self.overload)
        hi = np.float32(self.box + self.overload)
        return (np.full(3, lo, np.float32), np.full(3, hi, np.float32))


def plummer_mix(n: int, box: float, seed: int, n_clumps: int = 20,
                scale: float = 0.5, frac_clumped: float = 0.3):
    """70 % uniform, 30 % in Plummer clumps (scale radius ``scale``); centres
    uniform; radii r = a / sqrt(u^(-2/3) - 1), isotropic; wrapped mod L."""
    rng = np.random.default_rng(seed)
    n_uni = int(round((1.0 - frac_clumped) * n))
    n_cl = n - n_uni
    pos = np.empty((n, 3), np.float64)
    pos[:n_uni] = rng.random((n_uni, 3)) * box
    centres = rng.random((n_clumps, 3)) * box
    which = rng.integers(0, n_clumps, n_cl)
    u = rng.random(n_cl)
    u = np.clip(u, 1e-12, 1.0 - 1e-12)
    r = scale / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
    r = np.minimum(r, 0.45 * box)
    d = rng.normal(size=(n_cl, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    pos[n_uni:] = centres[which] + r[:, None] * d
    pos = np.mod(pos, box)
    pos = pos.astype(np.float32)
    pos[pos >= np.float32(box)] = np.float32(0.0)
    return pos[:, 0].copy(), pos[:, 1].copy(), pos[:, 2].copy()


def _nfw_radii(rng, n, c, r200):
    """Inverse-CDF sample of NFW radii truncated at r200."""
    def mu(x):
        return np.log1p(x) - x / (1.0 + x)
    mtot = mu(c)
    u = rng.random(n) * mtot
    # Newton iterations on mu(x) = u, x in (0, c]
    x = np.clip(u / mtot * c, 1e-6, c)
    for _ in range(60):
        f = mu(x) - u
        df = x / (1.0 + x) ** 2
        x = np.clip(x - f / np.maximum(df, 1e-30), 1e-9, c)
    return x * (r200 / c)


def clustered(n_side: int, box: float, n_halos: int, seed: int,
              rms_disp: float = 2.0, halo_min: float = 1e3,
              halo_max: float = 1e5, log=None):
    """Gaussian random field with P(k) ~ k^-2 (k from 2 pi / L to Nyquist),
    first-order Zel'dovich displacement scaled to RMS ``rms_disp``, plus
    ``n_halos`` NFW halos (c ~ U[5,15], N_h ~ logU[halo_min, halo_max],
    R200 = (3 N_h / (4 pi 200))^(1/3) at mean density 1).  Halo particles
    replace field particles so the total stays n_side^3.  ``log``, if given,
    receives a line after each stage."""
    say = log or (lambda msg: None)
    rng = np.random.default_rng(seed)
    n = n_side ** 3
    kf = 2.0 * np.pi / box
    k1 = np.fft.fftfreq(n_side, d=1.0 / n_side) * kf
    k1r = np.fft.rfftfreq(n_side, d=1.0 / n_side) * kf
    kx = k1[:, None, None]
    ky = k1[None, :, None]
    kz = k1r[None, None, :]
    k2 = kx * kx + ky * ky + kz * kz
    k2[0, 0, 0] = 1.0
    amp = 1.0 / np.sqrt(k2)  # |delta(k)| ~ k^-1, P ~ k^-2
    amp[0, 0, 0] = 0.0
    noise = rng.normal(size=(n_side, n_side, n_side)).astype(np.float32)
    dk = np.fft.rfftn(noise) * amp.astype(np.float32)
    del noise
    say(f"clustered {n_side}^3: density field done")
    disp = []
    for kk in (kx, ky, kz):
        psi_k = (1j * kk / k2).astype(np.complex64) * dk
        disp.append(np.fft.irfftn(psi_k, s=(n_side,) * 3, axes=(0, 1, 2)).astype(np.float32))
        del psi_k
        say(f"clustered {n_side}^3: displacement {len(disp)}/3 done")
    del dk
    rms = math.sqrt(sum(float(np.mean(d.astype(np.float64) ** 2)) for d in disp))
    scale = rms_disp / rms if rms > 0 else 0.0
    cell = box / n_side
    q = (np.arange(n_side, dtype=np.float32) + np.float32(0.5)) * np.float32(cell)
    pos = np.empty((n, 3), np.float32)
    pos[:, 0] = (q[:, None, None] + scale * disp[0]).reshape(-1)
    pos[:, 1] = (q[None, :, None] + scale * disp[1]).reshape(-1)
    pos[:, 2] = (q[None, None, :] + scale * disp[2]).reshape(-1)
    del disp

    # halos replace randomly chosen field particles
    lo, hi = math.log(halo_min), math.log(halo_max)
    sizes = np.exp(rng.uniform(lo, hi, n_halos)).astype(np.int64)
    total = int(sizes.sum())
    if total > 0.9 * n:
        sizes = (sizes * (0.9 * n / total)).astype(np.int64)
        total = int(sizes.sum())
    repl = rng.permutation(n)[:total]
    o = 0
    for h in range(n_halos):
        nh = int(sizes[h])
        if nh <= 0:
            continue
        c = rng.uniform(5.0, 15.0)
        r200 = (3.0 * nh / (4.0 * math.pi * 200.0)) ** (1.0 / 3.0)
        centre = rng.random(3) * box
        r = _nfw_radii(rng, nh, c, r200)
        d = rng.normal(size=(nh, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        pos[repl[o:o + nh]] = (centre + r[:, None] * d).astype(np.float32)
        o += nh
        if h % 250 == 249:
            say(f"clustered {n_side}^3: {h + 1}/{n_halos} halos placed")
    pos = np.mod(pos.astype(np.float64), box).astype(np.float32)
    pos[pos >= np.float32(box)] = np.float32(0.0)
    return pos[:, 0].copy(), pos[:, 1].copy(), pos[:, 2].copy()


def add_ghosts(x, y, z, box: float, overload: float) -> Load:
    """Append periodic images of every particle within ``overload`` of a face
    (faces, edges, corners in BasicDefinition.h:224-264 order), then shift
    everything by +overload (the reference's overload layout, Appendix of the
    domain model: alive region [o, L+o), tree box [0, L+2o))."""
    n = x.shape[0]
    L = np.float32(box)
    o = np.float32(overload)
    xs = [x, y, z]
    near_lo = [c < o for c in xs]
    near_hi = [c >= L - o for c in xs]
    parts_idx = [np.arange(n, dtype=np.int64)]
    shifts = [np.zeros((n, 3), np.float32)]
    for d in _NEIGHBORS:
        sel = np.ones(n, bool)
        for a in range(3):
            if d[a] == -1:
                sel &= near_lo[a]
            elif d[a] == 1:
                sel &= near_hi[a]
        idx = np.nonzero(sel)[0]
        if idx.size == 0:
            continue
        sh = np.zeros((idx.size, 3), np.float32)
        for a in range(3):
            # a particle near the low face images beyond the high face
            if d[a] == -1:
                sh[:, a] = L
            elif d[a] == 1:
                sh[:, a] = -L
        parts_idx.append(idx)
        shifts.append(sh)
    idx = np.concatenate(parts_idx)
    sh = np.concatenate(shifts)
    gx = (x[idx] + sh[:, 0]) + o
    gy = (y[idx] + sh[:, 1]) + o
    gz = (z[idx] + sh[:, 2]) + o
    return Load(gx.astype(np.float32), gy.astype(np.float32),
                gz.astype(np.float32), n, float(box), float(overload), idx)


def cell_sort(load: Load) -> Load:
    """Stable counting sort by unit cell as Particles::resortParticles does
    (Particles.cxx:402-488): key (floor(x)*ng + floor(y))*ng + floor(z) with
    ng = ceil(L + 2o); cells outside the grid go to the tail.  Returns the
    sorted load and the gather order."""
    ng = int(math.ceil(load.box + 2 * load.overload))
    ix = np.floor(load.x).astype(np.int64)
    iy = np.floor(load.y).astype(np.int64)
    iz = np.floor(load.z).astype(np.int64)
    inb = (ix >= 0) & (ix < ng) & (iy >= 0) & (iy < ng) & (iz >= 0) & (iz < ng)
    key = np.where(inb, (ix * ng + iy) * ng + iz, ng ** 3)
    order = np.argsort(key, kind="stable")
    return Load(load.x[order].copy(), load.y[order].copy(), load.z[order].copy(),
                load.n_alive, load.box, load.overload,
                load.alive_index[order].copy()), order


def config_load(name: str, seed: int | None = None, sort: bool = True,
                overload: float = 3.2, n_side: int | None = None, log=None):
    """Named loads.  'A': 64^3 Plummer mix in a 64 box (seed 1001);
    'B': 256^3 clustered + 500 halos (seed 1002); 'D': 128^3 clustered,
    halos scaled by volume (seed 1004); 'S<n>': small clustered n^3 test
    load in an n box."""
    if name == "A":
        s = 1001 if seed is None else seed
        ns = 64 if n_side is None else n_side
        x, y, z = plummer_mix(ns ** 3, float(ns), s)
        box = float(ns)
    elif name == "B":
        s = 1002 if seed is None else seed
        ns = 256 if n_side is None else n_side
        x, y, z = clustered(ns, float(ns), max(1, round(500 * (ns / 256) ** 3)), s, log=log)
        box = float(ns)
    elif name == "C":
        s = 1003 if seed is None else seed
        ns = 512 if n_side is None else n_side
        x, y, z = clustered(ns, float(ns), max(1, round(2000 * (ns / 512) ** 3)), s, log=log)
        box = float(ns)
    elif name == "D":
        s = 1004 if seed is None else seed
        ns = 128 if n_side is None else n_side
        x, y, z = clustered(ns, float(ns), max(1, round(500 * (ns / 256) ** 3)), s, log=log)
        box = float(ns)
    else:
        raise ValueError(f"unknown load {name!r}")
    load = add_ghosts(x, y, z, box, overload)
    if sort:
        load, _ = cell_sort(load)
    return load


def ensemble_loads(nboxes: int = 64, seed0: int = 1100, n_min: int = 32, n_max: int = 128,
                   overload: float = 3.2, log=None):
    """Independent periodic boxes: n_side ~ log-uniform in [n_min, n_max]
    (particle counts log-uniform between n_min^3 and n_max^3), box side
    n_side (mean density 1), the clustered generator with seed seed0 + b and
    halos scaled by volume (500 per 256^3); each box with its ghost shell and
    cell sort."""
    rng = np.random.default_rng(seed0)
    sides = np.rint(np.exp(rng.uniform(math.log(n_min), math.log(n_max), nboxes))).astype(int)
    loads = []
    for b, ns in enumerate(sides):
        x, y, z = clustered(int(ns), float(ns), max(1, round(500 * (ns / 256) ** 3)), seed0 + b)
        load, _ = cell_sort(add_ghosts(x, y, z, float(ns), overload))
        loads.append(load)
        if log and b % 16 == 15:
            log(f"ensemble: {b + 1}/{nboxes} boxes generated")
    return loads
