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
single steps, each one call or a short sequence of calls of the C ABI.
"""
from __future__ import annotations

import ctypes as C
import dataclasses

import numpy as np

from ._lib import Context, SubcycleParams

_libm = None


def _m():
    global _libm
    if _libm is None:
        _libm = C.CDLL("libm.so.6")
        _libm.powf.argtypes = [C.c_float, C.c_float]
        _libm.powf.restype = C.c_float
        _libm.atanf.argtypes = [C.c_float]
        _libm.atanf.restype = C.c_float
    return _libm


def map1_factor(pp: float, tau: float, adot: float, alpha: float) -> float:
    """prefactor * tau of Particles::map1 (Particles.cxx:746-747) in its
    float/double mix: powf(pp, 1 + 1/alpha) with the double exponent passed
    as float; 1.0 / (alpha * adot * pf) with a float product and a double
    quotient stored as float."""
    f = np.float32
    pf = f(_m().powf(f(pp), f(1.0 + 1.0 / float(f(alpha)))))
    prefactor = f(1.0 / float(f(alpha) * f(adot) * pf))
    return float(prefactor * f(tau))


def map2_coeff(gpscal: float, fscal: float, tau: float, step_fraction: float) -> float:
    """Kick coefficient of Particles::map2 (Particles.cxx:1230-1233)."""
    f = np.float32
    g = f(gpscal)
    divscal = g * g * g
    pi = f(4.0 * float(f(_m().atanf(f(1.0)))))
    return float(f(float(divscal) / 4.0 / float(pi) * fscal * tau * step_fraction))


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

    def _boxes(self):
        gmax = np.float32(max(self.ng))
        lo_t = np.zeros(3, np.float32)
        hi_t = np.full(3, gmax, np.float32)
        lo_f = np.full(3, np.float32(self.edge), np.float32)
        hi_f = np.array([np.float32(1.0 * g - float(np.float32(self.edge))) for g in self.ng],
                        np.float32)
        return lo_t, hi_t, lo_f, hi_f

    # Particles::map2 (RCB monopole branch)
    def map2(self, ts: TimeStepperValues, stepFraction: float = 1.0):
        c = map2_coeff(self.gpscal, ts.fscal, ts.tau, stepFraction)
        self.resortParticles()
        m = self.arrays["mass"]
        if isinstance(m, np.ndarray):
            m[:] = 1.0
        else:
            m.fill_(1.0)
        n = self.Np_last
        lo_t, hi_t, lo_f, hi_f = self._boxes()
        sl = {k: (None if v is None else v[:n]) for k, v in self.arrays.items()}
        self.ctx.build(sl["x"], sl["y"], sl["z"], sl["mass"], sl["vx"], sl["vy"], sl["vz"],
                       sl["phi"], sl["id"], sl["mask"], min_loc=lo_t, max_loc=hi_t,
                       device_ptrs=self.device_ptrs)
        self.ctx.lists(lo_f, hi_f)
        self.ctx.kick(c)

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
