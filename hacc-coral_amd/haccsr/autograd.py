"""Differentiable short-range force and sub-cycled kick-drift.

``ShortRangeForce`` is a ``torch.autograd.Function``: the forward pass
builds the tree and lists for the given positions and evaluates the leaf
kernel; the backward pass calls the hand-written adjoint (hsr_adjoint,
d(sum lambda . F)/dx with the topology frozen).  Only the positions of each
call are kept (a checkpoint); the backward pass rebuilds the tree and lists
from them, so memory stays at one tree per call in flight.

``kick_drift`` runs sub-steps v += c F(x); x += dt v (or the drift-kick-drift
form of Particles::subCycle, Particles.cxx:1176-1201), where c is the kick
coefficient map2 passes to the tree (Particles.cxx:1231-1234,
c = gpscal^3/(4 pi) * fscal * tau * stepFraction).  Periodic images are
re-derived from the box's own particles at every sub-step, so gradients
that reach an image flow back to the particle it images.
"""
from __future__ import annotations

import dataclasses

import numpy as np
import torch

from ._lib import Context


@dataclasses.dataclass
class Topology:
    """Tree and lists of one forward call (for frozen-topology checks)."""
    perm: np.ndarray        # tree position -> input index
    nodes: np.ndarray
    leaves: np.ndarray
    seg_ptr: np.ndarray
    seg_a: np.ndarray
    seg_b: np.ndarray
    entries: np.ndarray
    pos: np.ndarray | None = None  # the float positions the tree was built from (input order)


class ForceSolver:
    """A device context plus the box geometry of one rank's particles."""

    def __init__(self, tree_lo, tree_hi, force_lo, force_hi, nd=128,
                 fsrrmax=3.12, rsm=0.1, oa=0.1, record=False):
        self.ctx = Context(nd=nd, fsrrmax=fsrrmax, rsm=rsm, oa=oa)
        self.lo = np.asarray(tree_lo, np.float32)
        self.hi = np.asarray(tree_hi, np.float32)
        self.flo = np.asarray(force_lo, np.float32)
        self.fhi = np.asarray(force_hi, np.float32)
        self.record = record
        self.mass = None        # optional (n,) float32 cuda masses; ones if None
        self.topologies: list[Topology] = []
        self._pos = None        # the tensor the resident tree was built for
        self._ver = -1
        self._perm = None

    def _build(self, pos: torch.Tensor):
        """Tree + lists for pos (n x 3, float32, cuda); returns the
        permutation (tree position -> input index) as a device tensor."""
        n = pos.shape[0]
        p0 = self._pos
        if (p0 is not None and p0.data_ptr() == pos.data_ptr() and p0.shape == pos.shape
                and p0.stride() == pos.stride() and self._ver == pos._version):
            # same storage, unchanged since the resident tree was built (the
            # reference held in self._pos keeps the address from being reused)
            return self._perm
        if pos.dtype != torch.float32 or not pos.is_cuda or pos.dim() != 2 or pos.shape[1] != 3:
            raise ValueError("positions must be an (n, 3) float32 cuda tensor")
        x = pos[:, 0].contiguous().clone()
        y = pos[:, 1].contiguous().clone()
        z = pos[:, 2].contiguous().clone()
        if self.mass is None:
            m = torch.ones(n, dtype=torch.float32, device=pos.device)
        else:
            if self.mass.shape != (n,):
                raise ValueError("mass must have one entry per position")
            m = self.mass.to(torch.float32).contiguous().clone()
        ids = torch.arange(n, dtype=torch.int64, device=pos.device)
        torch.cuda.synchronize()
        self.ctx.build(x, y, z, m, ids=ids, min_loc=self.lo, max_loc=self.hi,
                       device_ptrs=True)
        self.ctx.lists(self.flo, self.fhi)
        self._pos, self._ver = pos, pos._version
        self._perm = ids
        return ids

    def forces(self, pos: torch.Tensor, record: bool | None = None) -> torch.Tensor:
        perm = self._build(pos)
        n = pos.shape[0]
        f = [torch.empty(n, dtype=torch.float32, device=pos.device) for _ in range(3)]
        self.ctx.forces(*f, device_ptrs=True)
        out = torch.empty((n, 3), dtype=torch.float32, device=pos.device)
        out[perm] = torch.stack(f, 1)
        if self.record if record is None else record:
            leaves, seg_ptr, sa, sb, ent = self.ctx.get_lists()
            self.topologies.append(Topology(perm.cpu().numpy(), self.ctx.nodes(), leaves,
                                            seg_ptr, sa, sb, ent, pos.detach().cpu().numpy()))
        return out

    def vjp(self, pos: torch.Tensor, lam: torch.Tensor) -> torch.Tensor:
        perm = self._build(pos)
        n = pos.shape[0]
        lt = lam.to(torch.float32)[perm]
        lx, ly, lz = (lt[:, k].contiguous() for k in range(3))
        g = [torch.empty(n, dtype=torch.float32, device=pos.device) for _ in range(3)]
        torch.cuda.synchronize()
        self.ctx.adjoint(lx, ly, lz, *g, device_ptrs=True)
        out = torch.empty((n, 3), dtype=torch.float32, device=pos.device)
        out[perm] = torch.stack(g, 1)
        return out


class ShortRangeForce(torch.autograd.Function):
    """F = ShortRangeForce.apply(pos, solver); pos (n x 3) float32 cuda."""

    @staticmethod
    def forward(ctx, pos, solver):
        ctx.solver = solver
        ctx.save_for_backward(pos)
        return solver.forces(pos.detach())

    @staticmethod
    def backward(ctx, gF):
        (pos,) = ctx.saved_tensors
        return ctx.solver.vjp(pos, gF.contiguous()), None


@dataclasses.dataclass
class Images:
    """Periodic images: entry k of the rank's array is alive particle
    index[k] shifted by shift[k] (the first n_alive entries are the alive
    particles themselves)."""
    index: torch.Tensor     # int64 (n_all,)
    shift: torch.Tensor     # float32 (n_all, 3)
    n_alive: int


def images_from_layout(alive_x, alive_y, alive_z, box: float, overload: float,
                       device="cuda") -> Images:
    """Image map of the ghost layout of haccsr.ics.add_ghosts (+overload
    shift included)."""
    from . import ics
    load = ics.add_ghosts(alive_x, alive_y, alive_z, box, overload)
    idx = torch.from_numpy(load.alive_index).to(device)
    base = np.stack([alive_x, alive_y, alive_z], 1)[load.alive_index]
    full = np.stack([load.x, load.y, load.z], 1)
    shift = torch.from_numpy((full.astype(np.float64) - base.astype(np.float64)).astype(np.float32)).to(device)
    return Images(idx, shift, int(load.n_alive))


def kick_drift(x: torch.Tensor, v: torch.Tensor, solver: ForceSolver, nsub: int,
               dt: float, kick: float, images: Images | None = None,
               scheme: str = "kd"):
    """nsub differentiable sub-steps through ShortRangeForce.  x, v: the
    box's own particles (n_alive x 3).

    scheme "kd":  v += kick * F(x); x += dt * v.
    scheme "dkd": x += dt/2 * v; v += kick * F(x); x += dt/2 * v, the shape of
                  Particles::subCycle (half stream, kick, half stream)."""
    if scheme not in ("kd", "dkd"):
        raise ValueError("scheme must be 'kd' or 'dkd'")

    def force(x):
        pos = x if images is None else x[images.index] + images.shift
        F = ShortRangeForce.apply(pos, solver)
        return F if images is None else F[: images.n_alive]

    for _ in range(nsub):
        if scheme == "kd":
            v = v + kick * force(x)
            x = x + dt * v
        else:
            x = x + (0.5 * dt) * v
            v = v + kick * force(x)
            x = x + (0.5 * dt) * v
    return x, v
