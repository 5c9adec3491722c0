"""The sub-cycle around the short-range kick, on the device.

``Particles`` mirrors the methods of the reference's Particles class that
drive the RCB force tree (Particles.h:160-175): ``map1`` (drift,
Particles.cxx:732-758), ``resortParticles`` (Particles.cxx:402-488),
``map2`` (resort, unit masses, RCB monopole tree over the particles inside
the grid and its kick, Particles.cxx:1205-1280) and ``subCycle``
(Particles.cxx:1176-1201).  The particle arrays stay with the caller: numpy
arrays (host memory, staged by the library) or float32/int64/int16 device
tensors (used in place).  ``subCycle`` is one hsr_subcycle call that keeps the
particles on the device for all nsub sub-steps; ``map1`` and ``map2`` are the
single steps (hsr_drift, hsr_map2), one C ABI call each.
``refreshParticles`` is the one-rank overload refresh (hsr_refresh).
"""
from __future__ import annotations

import dataclasses

import numpy as np

from ._lib import CapacityError, Context, SubcycleParams, domain

@dataclasses.dataclass
class TimeStepperValues:
    """The TimeStepper accessors map1 and map2 read (TimeStepper.h:20-36)."""
    pp: float
    tau: float
    tau2: float
    adot: float
    fscal: float


class Particles:
    """Sub-cycle methods of Particles over libhaccsr (RCB monopole tree)."""

    ARRAYS = ("x", "y", "z", "vx", "vy", "vz", "mass", "phi", "id", "mask")

    def __init__(self, ng, *, alpha: float, gpscal: float, edge: float, nsub: int = 5,
                 fsrrmax: float = 3.12, rsm: float = 0.1, oa: float = 0.1,
                 rcb_tree_ppn: int = 128, skip_stream: bool = False,
                 skip_kick: bool = False, device: int = 0):
        self.ng = [int(g) for g in ng]
        self.alpha, self.gpscal, self.edge = float(alpha), float(gpscal), float(edge)
        self.nsub = int(nsub)
        self.skip_stream, self.skip_kick = bool(skip_stream), bool(skip_kick)
        self.ctx = Context(nd=rcb_tree_ppn, fsrrmax=fsrrmax, rsm=rsm, oa=oa, device=device)
        self.Np_last = -1
        self.arrays = {}
        self.device_ptrs = False

    def close(self):
        self.ctx.close()

    def set_particles(self, x, y, z, vx, vy, vz, mass, phi=None, ids=None, mask=None):
        """Hand over the particle arrays (all numpy, or all device tensors)."""
        self.arrays = dict(x=x, y=y, z=z, vx=vx, vy=vy, vz=vz, mass=mass, phi=phi,
                           id=ids, mask=mask)
        self.device_ptrs = not isinstance(x, np.ndarray)

    @property
    def Np_local_total(self) -> int:
        return int(self.arrays["x"].shape[0])

    def _a(self, *names):
        return [self.arrays[k] for k in names]

    # Particles::map1
    def map1(self, pp: float, tau: float, adot: float):
        x, y, z, vx, vy, vz = self._a("x", "y", "z", "vx", "vy", "vz")
        self.ctx.drift(x, y, z, vx, vy, vz, pp, tau, adot, self.alpha,
                       device_ptrs=self.device_ptrs)

    # Particles::resortParticles
    def resortParticles(self):
        a = self.arrays
        self.Np_last = self.ctx.resort(self.ng, a["x"], a["y"], a["z"], a["vx"], a["vy"],
                                       a["vz"], a["mass"], a["phi"], a["id"], a["mask"],
                                       device_ptrs=self.device_ptrs)

    # Particles::map2 (RCB monopole branch)
    def map2(self, ts: TimeStepperValues, stepFraction: float = 1.0):
        a = self.arrays
        self.Np_last = self.ctx.map2(self.params(ts), stepFraction, a["x"], a["y"], a["z"],
                                     a["vx"], a["vy"], a["vz"], a["mass"], a["phi"], a["id"],
                                     a["mask"], device_ptrs=self.device_ptrs)

    def params(self, ts: TimeStepperValues) -> SubcycleParams:
        p = SubcycleParams()
        p.nsub = self.nsub
        p.skip_stream = int(self.skip_stream)
        p.skip_kick = int(self.skip_kick)
        for d in range(3):
            p.ng[d] = self.ng[d]
        p.alpha, p.gpscal, p.edge = self.alpha, self.gpscal, self.edge
        p.pp, p.tau, p.tau2, p.adot, p.fscal = ts.pp, ts.tau, ts.tau2, ts.adot, ts.fscal
        return p

    # Particles::subCycle: one device-resident call
    def subCycle(self, ts: TimeStepperValues):
        a = self.arrays
        self.Np_last = self.ctx.subcycle(self.params(ts), a["x"], a["y"], a["z"], a["vx"],
                                         a["vy"], a["vz"], a["mass"], a["phi"], a["id"],
                                         a["mask"], device_ptrs=self.device_ptrs)

    # the same cycle as single steps (map1, map2, map1 per sub-step)
    def subCycle_steps(self, ts: TimeStepperValues, after_step=None):
        sf = 1.0 / self.nsub
        for step in range(self.nsub):
            if not self.skip_stream:
                self.map1(ts.pp, np.float32(sf * ts.tau2), ts.adot)
            if not self.skip_kick:
                self.map2(ts, sf)
            if not self.skip_stream:
                self.map1(ts.pp, np.float32(sf * ts.tau2), ts.adot)
            if after_step is not None:
                after_step(step)

    # MC3Extras::particleExtras' refresh on one rank (MC3Extras.cxx:330-338,
    # 660-700): copyAliveIntoVectors, ParticleExchange, copyAllFromVectors,
    # resortParticles
    def refreshParticles(self, rL: float, oL: float, ng: int, anow: float,
                         capacity: int | None = None) -> int:
        """Refresh the overload copies for the domain of the deck's rL, oL and
        ng.  The particle arrays are replaced by arrays of the refreshed
        length (views of new buffers of `capacity`, grown if the exchange
        needs more); returns the new particle count."""
        dom = domain(rL, oL, ng)
        if any(g != dom.ng_local_total for g in self.ng):
            raise ValueError(f"refreshParticles: the grid {self.ng} is not the domain's "
                             f"ng_local_total {dom.ng_local_total}")
        n = self.Np_local_total
        cap = max(int(capacity) if capacity else n + n // 4 + 1024, n)
        for attempt in range(2):
            bufs = {k: self._grown(k, cap) for k in self.ARRAYS}
            try:
                count, inside = self.ctx.refresh(dom, anow, n, *(bufs[k] for k in self.ARRAYS),
                                                 device_ptrs=self.device_ptrs)
                break
            except CapacityError as e:
                if attempt:
                    raise
                cap = e.required
        self.arrays = {k: b[:count] for k, b in bufs.items()}
        self.Np_last = inside
        return count

    def _grown(self, name: str, cap: int):
        a = self.arrays[name]
        n = int(a.shape[0])
        if isinstance(a, np.ndarray):
            b = np.empty(cap, a.dtype)
        else:
            b = a.new_empty(cap)
        b[:n] = a
        return b
