"""Python mirror of the reference's constructor surface for this path.

RCBForceTree<TDPTS>(minLoc, maxLoc, minForceLoc, maxForceLoc, count,
                    xLoc, yLoc, zLoc, xVel, yVel, zVel, mass, phiLoc, idLoc,
                    maskLoc, avgMass, fsm, r, oa, nd=1, ds=1, tmin=128,
                    fl=0, fcoeff=0.0, ppc=0.9)      (RCBForceTree.h:96-124)

builds the tree (permuting all ten arrays in place), builds the interaction
lists of the leaves inside the force box and applies the short-range kick
v += fcoeff * F -- all inside the constructor (RCBForceTree.cxx:336-450).
The ForceLaw hierarchy (ForceLaw.h:10-127) keeps its names and
constructors; on the device a law is an evaluator: ForceLawSR with
FGridEvalPoly (orders 5, 6), FGridEvalFit or FGridEvalInterp.
ForceLawNewton raises (no CPU fallback).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib


class FGrid:
    """FGrid (ForceLaw.cxx:23-33): only rmax() matters to the device path."""

    def rmax(self) -> float:
        return float(np.float32(3.116326355))


class FGridEval:
    pass


class FGridEvalFit(FGridEval):
    """FGridEvalFit (ForceLaw.cxx:71-81): the fitted grid force FGrid::fgor,
    evaluated per pair by the device kernel (LAW_SR_FIT)."""

    def __init__(self, fg: FGrid):
        self.fg = fg


class FGridEvalInterp(FGridEval):
    """FGridEvalInterp(FGrid*, nInterp) (ForceLaw.cxx:145-171): FGrid::fgor
    tabulated on nInterp points of r^2 (FGrid::fgor_r2_interp, computed by
    the C++ drop-in library libhaccsr_cxx.so) and interpolated linearly by the
    device kernel (LAW_SR_INTERP)."""

    def __init__(self, fg: FGrid, nInterp: int):
        if int(nInterp) < 2:
            raise ValueError("FGridEvalInterp: nInterp must be >= 2")
        self.fg = fg
        self.n = int(nInterp)
        self._r2 = np.zeros(self.n, np.float32)
        self._f = np.zeros(self.n, np.float32)
        lib = ctypes.CDLL(_lib.CXX_LIB_PATH)
        lib.hsr_cxx_fgrid_table.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        lib.hsr_cxx_fgrid_table.restype = None
        lib.hsr_cxx_fgrid_table(self.n, self._r2.ctypes.data, self._f.ctypes.data)

    def nInterp(self) -> int:
        return self.n

    def table(self):
        return self._r2, self._f


class FGridEvalPoly(FGridEval):
    """FGridEvalPoly (ForceLaw.cxx:93-141).  The reference selects the
    polynomial at compile time with POLY_ORDER (stock build: 6); here it is
    a constructor argument."""

    def __init__(self, fg: FGrid, order: int = 6):
        if order not in (5, 6):
            raise ValueError("FGridEvalPoly: order must be 5 or 6")
        self.fg = fg
        self.order = order


class ForceLaw:
    pass


class ForceLawNewton(ForceLaw):
    """ForceLawNewton (ForceLaw.h:105-110)."""


class ForceLawSR(ForceLaw):
    """ForceLawSR(FGridEval*, rsm) (ForceLaw.h:114-127)."""

    def __init__(self, fgore: FGridEval, rsm: float):
        self.fgore = fgore
        self.rsm = float(rsm)


def _law_id(fl, kernel: str) -> int:
    if kernel == "step16":
        # BG/Q production build: nbody1 always calls Step16_int with the
        # tree's rsm and fsrrmax, whatever ForceLaw was passed
        # (RCBForceTree.cxx:585-597)
        return _lib.LAW_STEP16
    if fl is None or isinstance(fl, ForceLawNewton):
        raise _lib.HsrError("ForceLawNewton is not implemented by the device kernel")
    if isinstance(fl, ForceLawSR) and isinstance(fl.fgore, FGridEvalPoly):
        return _lib.LAW_SR_POLY5 if fl.fgore.order == 5 else _lib.LAW_SR_POLY6
    if isinstance(fl, ForceLawSR) and isinstance(fl.fgore, FGridEvalFit):
        return _lib.LAW_SR_FIT
    if isinstance(fl, ForceLawSR) and isinstance(fl.fgore, FGridEvalInterp):
        return _lib.LAW_SR_INTERP
    raise _lib.HsrError(f"{type(fl).__name__}/{type(getattr(fl, 'fgore', None)).__name__}"
                        " is not implemented by the device kernel")


class RCBForceTree:
    """Drop-in for RCBForceTree<TDPTS>.  Arrays are numpy float32 (ids int64,
    mask uint16), modified in place like the reference's caller-owned
    arrays.  ``kernel`` selects the BG/Q Step16 law (default) or the
    generic ForceLaw kernel."""

    TDPTS = 1

    def __init__(self, minLoc, maxLoc, minForceLoc, maxForceLoc, count, xLoc,
                 yLoc, zLoc, xVel, yVel, zVel, mass, phiLoc, idLoc, maskLoc,
                 avgMass, fsm, r, oa, nd=1, ds=1, tmin=128, fl=None,
                 fcoeff=0.0, ppc=0.9, kernel: str = "step16", device: int = 0):
        if fl is None:
            fcoeff = 1.0  # RCBForceTree.cxx:399-404
        law = _law_id(fl, kernel)
        rsm = float(r)
        if isinstance(fl, ForceLawSR) and kernel != "step16":
            rsm = fl.rsm
        table = fl.fgore.table() if law == _lib.LAW_SR_INTERP else None
        self.ctx = _lib.Context(nd=nd, fsrrmax=fsm, rsm=rsm, oa=oa, ppc=ppc,
                                law=law, device=device, ds=ds, tmin=tmin,
                                tdpts=self.TDPTS, interp_table=table)
        arrs = [xLoc, yLoc, zLoc, xVel, yVel, zVel, mass, phiLoc]
        for a in arrs:
            if a.dtype != np.float32 or not a.flags["C_CONTIGUOUS"]:
                raise TypeError("particle arrays must be contiguous float32")
        n = int(count)
        self.ctx.build(xLoc[:n], yLoc[:n], zLoc[:n], mass[:n], xVel[:n], yVel[:n],
                       zVel[:n], phiLoc[:n], idLoc[:n], maskLoc[:n],
                       min_loc=minLoc, max_loc=maxLoc)
        self.ctx.lists(minForceLoc, maxForceLoc)
        self.ctx.kick(float(fcoeff))
        self.stats = self.ctx.stats()

    def printStats(self):
        s = self.stats
        mean = 0.0
        print("\ttree post-build statistics (local for rank 0):")
        print(f"\t\tparticles: {float(s.count):.2f}")
        print(f"\t\tnodes: {float(s.nodes):.2f}")
        print(f"\t\tleaves: {float(s.leaves + s.empty_leaves):.2f} (empty: {float(s.empty_leaves):.2f})")
        if s.leaves:
            mean = self.ctx.nodes()["count"][1:][(self.ctx.nodes()["cl"][1:] == 0)].sum() / s.leaves
        print(f"\t\tmean ppn: {mean:.2f} (max ppn: {s.max_ppn})")
        print(f"\t\tbuild time: {s.build_ms * 1e-3:g} s")


class RCBMonopoleForceTree(RCBForceTree):
    TDPTS = 1


class RCBQuadrupoleForceTree(RCBForceTree):
    TDPTS = 12
