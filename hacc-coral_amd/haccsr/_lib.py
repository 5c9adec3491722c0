"""ctypes binding of libhaccsr.so (include/hacc_sr.h).

The library is built in-tree (hacc-coral_amd/libhaccsr.so) by
``make -C hacc-coral_amd``.  There is no CPU
fallback: every entry point runs on a HIP device and raises when the library
or the device is missing.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# HSR_LIBRARY overrides the library path (e.g. a build with other kernel
# shapes, make KDEFS=...)
LIB_PATH = os.environ.get("HSR_LIBRARY") or os.path.join(PKG_ROOT, "libhaccsr.so")
CXX_LIB_PATH = os.path.join(PKG_ROOT, "libhaccsr_cxx.so")  # C++ drop-ins (ForceLaw, ...)

HSR_HOST_PTRS, HSR_DEVICE_PTRS = 0, 1
LAW_STEP16, LAW_SR_POLY5, LAW_SR_POLY6, LAW_NEWTON = 0, 1, 2, 3
LAW_SR_FIT, LAW_SR_INTERP = 4, 5

# hsr_status codes (include/hacc_sr.h)
HSR_OK, HSR_ERR_ARG, HSR_ERR_HIP, HSR_ERR_UNSUPPORTED = 0, 1, 2, 3
HSR_ERR_STATE, HSR_ERR_NOMEM, HSR_ERR_NUMERIC = 4, 5, 6

EXPORTS = [
    "hsr_last_error", "hsr_version", "hsr_create", "hsr_destroy",
    "hsr_get_stream", "hsr_build", "hsr_lists", "hsr_forces", "hsr_kick",
    "hsr_adjoint", "hsr_get_stats", "hsr_get_nodes", "hsr_get_perm",
    "hsr_get_lists", "hsr_build_many", "hsr_lists_many", "hsr_set_tuning",
    "hsr_get_tuning", "hsr_forces_f64", "hsr_leaf_momentum", "hsr_drift",
    "hsr_resort", "hsr_map2", "hsr_subcycle", "hsr_get_pseudo_masses", "hsr_cic",
    "hsr_inverse_cic", "hsr_domain_init", "hsr_refresh", "hsr_resort_noalloc",
    "hsr_set_law_table", "hsr_forces_f64_at", "hsr_adjoint_f64_at",
]


class HsrError(RuntimeError):
    pass


class CapacityError(HsrError):
    """hsr_refresh needs room for `required` particles."""

    def __init__(self, msg, required):
        super().__init__(msg)
        self.required = required


class Params(C.Structure):
    _fields_ = [("device", C.c_int), ("tdpts", C.c_int), ("law", C.c_int),
                ("n_interp", C.c_int), ("nd", C.c_int64), ("ds", C.c_int64),
                ("tmin", C.c_int64), ("fsrrmax", C.c_float), ("rsm", C.c_float),
                ("oa", C.c_float), ("tan_oa", C.c_float), ("ppc", C.c_float),
                ("reserved1", C.c_float * 3)]


class Node(C.Structure):
    _fields_ = [("count", C.c_int64), ("offset", C.c_int64),
                ("cl", C.c_int64), ("cr", C.c_int64),
                ("ppm", C.c_float * 1), ("tdr", C.c_float),
                ("xmin", C.c_float * 3), ("xmax", C.c_float * 3),
                ("xc", C.c_float * 3)]


NODE_DTYPE = np.dtype([("count", "<i8"), ("offset", "<i8"), ("cl", "<i8"),
                       ("cr", "<i8"), ("ppm", "<f4", (1,)), ("tdr", "<f4"),
                       ("xmin", "<f4", (3,)), ("xmax", "<f4", (3,)),
                       ("xc", "<f4", (3,))], align=True)
assert NODE_DTYPE.itemsize == C.sizeof(Node)


class Stats(C.Structure):
    _fields_ = [("count", C.c_int64), ("nodes", C.c_int64), ("leaves", C.c_int64),
                ("empty_leaves", C.c_int64), ("max_ppn", C.c_int64),
                ("listed_leaves", C.c_int64), ("segments", C.c_int64),
                ("entries", C.c_int64), ("pairs", C.c_int64),
                ("max_depth", C.c_int64), ("build_ms", C.c_double),
                ("lists_ms", C.c_double), ("force_ms", C.c_double),
                ("kernel_ms", C.c_double), ("adjoint_ms", C.c_double),
                ("evaluated_pairs", C.c_int64), ("drift_ms", C.c_double),
                ("resort_ms", C.c_double), ("subcycle_ms", C.c_double),
                ("drift_factor", C.c_double), ("kick_coeff", C.c_double),
                ("cic_ms", C.c_double), ("inverse_cic_ms", C.c_double),
                ("refresh_ms", C.c_double), ("refresh_alive", C.c_int64),
                ("refresh_dead", C.c_int64), ("unmasked_pairs", C.c_int64),
                ("graph_captures", C.c_int64), ("graph_replays", C.c_int64)]


class Domain(C.Structure):
    """hsr_domain: a one-rank Domain (Domain.cxx:34-84) from the deck's rL,
    oL and ng; the other fields are derived by hsr_domain_init."""
    _fields_ = [("rL", C.c_float), ("oL_in", C.c_float), ("ng", C.c_int32),
                ("ng_overload", C.c_int32), ("ng_local_total", C.c_int32),
                ("oL", C.c_float), ("phys2grid_pos", C.c_float),
                ("grid2phys_pos", C.c_float), ("phys2grid_vel", C.c_float),
                ("grid2phys_vel", C.c_float)]


class SubcycleParams(C.Structure):
    """hsr_subcycle_params: Particles' sub-cycle settings and the
    TimeStepper values map1/map2 read (Particles.cxx:1176-1233)."""
    _fields_ = [("nsub", C.c_int), ("skip_stream", C.c_int), ("skip_kick", C.c_int),
                ("ng", C.c_int32 * 3), ("alpha", C.c_float), ("gpscal", C.c_float),
                ("edge", C.c_float), ("reserved0", C.c_float), ("pp", C.c_double),
                ("tau", C.c_double), ("tau2", C.c_double), ("adot", C.c_double),
                ("fscal", C.c_double)]


class Tuning(C.Structure):
    _fields_ = [("i_per_lane", C.c_int), ("tile", C.c_int), ("chunk", C.c_int),
                ("waves_per_cu", C.c_int), ("use_graph", C.c_int),
                ("adjoint_mode", C.c_int), ("cull_off", C.c_int),
                ("tree_order_targets", C.c_int), ("force_pass", C.c_int),
                ("max_plan_bytes", C.c_int64), ("unroll", C.c_int), ("chain_hops", C.c_int)]


_lib = None


def load(path: str | None = None):
    """Load libhaccsr.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or os.environ.get("HSR_LIBRARY") or LIB_PATH
    if not os.path.exists(p):
        raise HsrError(f"{p} is missing: build it with `make -C hacc-coral_amd`")
    # PyTorch ships its own HIP runtime.  Loaded first, its symbols are the
    # ones libhaccsr.so binds to, so the process has one runtime and device
    # tensors' pointers are valid here; loaded after the library, torch finds
    # no device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(p)
    vp, i64, f32p = C.c_void_p, C.c_int64, C.c_void_p
    L.hsr_last_error.restype = C.c_char_p
    L.hsr_version.restype = C.c_char_p
    L.hsr_create.argtypes = [C.POINTER(C.c_void_p), C.POINTER(Params)]
    L.hsr_destroy.argtypes = [vp]
    L.hsr_get_stream.argtypes = [vp, C.POINTER(C.c_void_p)]
    L.hsr_build.argtypes = [vp, i64] + [vp] * 10 + [vp, vp, C.c_int]
    L.hsr_build_many.argtypes = [vp, i64, vp] + [vp] * 10 + [vp, vp, C.c_int]
    L.hsr_lists.argtypes = [vp, vp, vp]
    L.hsr_lists_many.argtypes = [vp, vp, vp]
    L.hsr_forces.argtypes = [vp, vp, vp, vp, C.c_int]
    L.hsr_kick.argtypes = [vp, C.c_float]
    L.hsr_adjoint.argtypes = [vp] + [vp] * 6 + [C.c_int]
    L.hsr_forces_f64.argtypes = [vp] + [vp] * 4 + [C.c_int]
    L.hsr_leaf_momentum.argtypes = [vp] + [vp] * 4 + [C.c_int]
    L.hsr_drift.argtypes = [vp, i64] + [vp] * 6 + [C.c_float] * 4 + [C.c_int]
    L.hsr_resort.argtypes = [vp, i64, vp] + [vp] * 10 + [vp, C.c_int]
    L.hsr_resort_noalloc.argtypes = [vp, i64, vp] + [vp] * 10 + [vp, C.c_int]
    L.hsr_set_law_table.argtypes = [vp, vp, vp, i64]
    L.hsr_forces_f64_at.argtypes = [vp] + [vp] * 6
    L.hsr_adjoint_f64_at.argtypes = [vp] + [vp] * 9
    L.hsr_map2.argtypes = [vp, C.POINTER(SubcycleParams), C.c_double, i64] + [vp] * 10 + \
        [vp, C.c_int]
    L.hsr_subcycle.argtypes = [vp, C.POINTER(SubcycleParams), i64] + [vp] * 10 + [vp, C.c_int]
    L.hsr_cic.argtypes = [vp, i64, vp, vp, vp, vp, C.c_float, vp, C.c_int]
    L.hsr_inverse_cic.argtypes = [vp, i64, vp, vp, vp, vp, vp, C.c_float, C.c_float, vp,
                                  C.c_int]
    L.hsr_domain_init.argtypes = [C.POINTER(Domain)]
    L.hsr_refresh.argtypes = [vp, C.POINTER(Domain), C.c_float, C.POINTER(C.c_int64), i64] + \
        [vp] * 10 + [vp, C.c_int]
    L.hsr_get_stats.argtypes = [vp, C.POINTER(Stats)]
    L.hsr_get_nodes.argtypes = [vp, vp, i64]
    L.hsr_get_pseudo_masses.argtypes = [vp, vp, i64]
    L.hsr_get_perm.argtypes = [vp, vp, i64]
    L.hsr_get_lists.argtypes = [vp, vp, vp, vp, vp, vp, i64, i64]
    L.hsr_set_tuning.argtypes = [vp, C.POINTER(Tuning)]
    L.hsr_get_tuning.argtypes = [vp, C.POINTER(Tuning)]
    for name in EXPORTS:
        if name not in ("hsr_last_error", "hsr_version"):
            getattr(L, name).restype = C.c_int
    _lib = L
    return L


def check(status: int, what: str):
    if status != 0:
        msg = load().hsr_last_error().decode(errors="replace")
        raise HsrError(f"{what} failed (status {status}): {msg}")


def tanf(x: float) -> float:
    """tanf(oa) as the host libm evaluates it (RCBForceTree.cxx:381)."""
    libm = C.CDLL("libm.so.6")
    libm.tanf.argtypes = [C.c_float]
    libm.tanf.restype = C.c_float
    return float(libm.tanf(C.c_float(x)))


def _ptr(a):
    """Address of a numpy array or torch tensor (None -> NULL)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        assert a.flags["C_CONTIGUOUS"]
        return a.ctypes.data
    return int(a.data_ptr())


class Context:
    """One device context (hsr_create ... hsr_destroy)."""

    def __init__(self, nd: int = 128, fsrrmax: float = 3.12, rsm: float = 0.1,
                 oa: float = 0.1, ppc: float = 0.9, law: int = LAW_STEP16,
                 device: int = 0, ds: int = 2, tmin: int = 128, tdpts: int = 1,
                 interp_table=None):
        """interp_table: (r2, f) of an FGridEvalInterp for LAW_SR_INTERP
        (forcetree.FGridEvalInterp(...).table())."""
        self.L = load()
        p = Params()
        p.device = device
        p.tdpts = tdpts
        p.law = law
        if law == LAW_SR_INTERP:
            if interp_table is None:
                raise HsrError("LAW_SR_INTERP needs interp_table=(r2, f)")
            r2t = np.ascontiguousarray(interp_table[0], np.float32)
            ft = np.ascontiguousarray(interp_table[1], np.float32)
            p.n_interp = int(r2t.shape[0])
        p.nd = int(nd)
        p.ds = int(ds)
        p.tmin = int(tmin)
        p.fsrrmax = float(np.float32(fsrrmax))
        p.rsm = float(np.float32(rsm))
        p.oa = float(np.float32(oa))
        p.tan_oa = tanf(float(np.float32(oa)))
        p.ppc = float(np.float32(ppc))
        self.params = p
        h = C.c_void_p()
        check(self.L.hsr_create(C.byref(h), C.byref(p)), "hsr_create")
        self.h = h
        self.n = 0
        if law == LAW_SR_INTERP:
            check(self.L.hsr_set_law_table(self.h, r2t.ctypes.data, ft.ctypes.data,
                                           r2t.shape[0]), "hsr_set_law_table")

    def close(self):
        if getattr(self, "h", None):
            self.L.hsr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- phases
    def build(self, x, y, z, mass, vx=None, vy=None, vz=None, phi=None,
              ids=None, mask=None, min_loc=(0, 0, 0), max_loc=(1, 1, 1),
              device_ptrs=False):
        n = int(x.shape[0])
        lo = np.ascontiguousarray(min_loc, np.float32)
        hi = np.ascontiguousarray(max_loc, np.float32)
        fl = HSR_DEVICE_PTRS if device_ptrs else HSR_HOST_PTRS
        check(self.L.hsr_build(self.h, n, _ptr(x), _ptr(y), _ptr(z), _ptr(vx),
                               _ptr(vy), _ptr(vz), _ptr(mass), _ptr(phi),
                               _ptr(ids), _ptr(mask), lo.ctypes.data,
                               hi.ctypes.data, fl), "hsr_build")
        self.n = n

    def build_many(self, offsets, x, y, z, mass, min_loc, max_loc, vx=None,
                   vy=None, vz=None, phi=None, ids=None, mask=None,
                   device_ptrs=False):
        off = np.ascontiguousarray(offsets, np.int64)
        lo = np.ascontiguousarray(min_loc, np.float32).reshape(-1)
        hi = np.ascontiguousarray(max_loc, np.float32).reshape(-1)
        fl = HSR_DEVICE_PTRS if device_ptrs else HSR_HOST_PTRS
        check(self.L.hsr_build_many(self.h, off.shape[0] - 1, off.ctypes.data,
                                    _ptr(x), _ptr(y), _ptr(z), _ptr(vx), _ptr(vy),
                                    _ptr(vz), _ptr(mass), _ptr(phi), _ptr(ids),
                                    _ptr(mask), lo.ctypes.data, hi.ctypes.data, fl),
              "hsr_build_many")
        self.n = int(off[-1])

    def lists(self, min_force, max_force):
        lo = np.ascontiguousarray(min_force, np.float32)
        hi = np.ascontiguousarray(max_force, np.float32)
        check(self.L.hsr_lists(self.h, lo.ctypes.data, hi.ctypes.data), "hsr_lists")

    def lists_many(self, min_force, max_force):
        lo = np.ascontiguousarray(min_force, np.float32).reshape(-1)
        hi = np.ascontiguousarray(max_force, np.float32).reshape(-1)
        check(self.L.hsr_lists_many(self.h, lo.ctypes.data, hi.ctypes.data),
              "hsr_lists_many")

    def forces(self, ax=None, ay=None, az=None, device_ptrs=False):
        if ax is None:
            ax = np.zeros(self.n, np.float32)
            ay = np.zeros(self.n, np.float32)
            az = np.zeros(self.n, np.float32)
        fl = HSR_DEVICE_PTRS if device_ptrs else HSR_HOST_PTRS
        check(self.L.hsr_forces(self.h, _ptr(ax), _ptr(ay), _ptr(az), fl), "hsr_forces")
        return ax, ay, az

    def kick(self, fcoeff: float):
        check(self.L.hsr_kick(self.h, C.c_float(fcoeff)), "hsr_kick")

    def adjoint(self, lx, ly, lz, gx=None, gy=None, gz=None, device_ptrs=False):
        """g = d(sum lambda . F)/dx (tree order).  Host arrays by default."""
        if gx is None:
            gx = np.zeros(self.n, np.float32)
            gy = np.zeros(self.n, np.float32)
            gz = np.zeros(self.n, np.float32)
        fl = HSR_DEVICE_PTRS if device_ptrs else HSR_HOST_PTRS
        check(self.L.hsr_adjoint(self.h, _ptr(lx), _ptr(ly), _ptr(lz), _ptr(gx),
                                 _ptr(gy), _ptr(gz), fl), "hsr_adjoint")
        return gx, gy, gz

    def forces_f64(self, ax=None, ay=None, az=None, leaf_mom=None, device_ptrs=False,
                   with_momentum=False):
        """FP64 evaluation of the same lists (tree order).  Host arrays by
        default; with_momentum also returns sum_i m_i a_i per listed leaf."""
        if ax is None:
            ax, ay, az = (np.zeros(self.n) for _ in range(3))
        if with_momentum and leaf_mom is None:
            leaf_mom = np.zeros((max(self.stats().listed_leaves, 1), 3))
        fl = HSR_DEVICE_PTRS if device_ptrs else HSR_HOST_PTRS
        check(self.L.hsr_forces_f64(self.h, _ptr(ax), _ptr(ay), _ptr(az), _ptr(leaf_mom), fl),
              "hsr_forces_f64")
        return (ax, ay, az, leaf_mom) if with_momentum else (ax, ay, az)

    def forces_f64_at(self, x, y, z):
        """FP64 forces at double positions (tree order, cuda float64
        tensors) with the current tree and lists frozen; returns (ax, ay, az)
        as cuda float64 tensors."""
        import torch
        out = [torch.empty_like(x) for _ in range(3)]
        check(self.L.hsr_forces_f64_at(self.h, _ptr(x), _ptr(y), _ptr(z), *map(_ptr, out)),
              "hsr_forces_f64_at")
        return tuple(out)

    def adjoint_f64_at(self, x, y, z, lx, ly, lz):
        """FP64 d(sum lambda . F)/dx at double positions (tree order, cuda
        float64 tensors), topology frozen; returns (gx, gy, gz)."""
        import torch
        out = [torch.empty_like(x) for _ in range(3)]
        check(self.L.hsr_adjoint_f64_at(self.h, _ptr(x), _ptr(y), _ptr(z), _ptr(lx), _ptr(ly),
                                        _ptr(lz), *map(_ptr, out)), "hsr_adjoint_f64_at")
        return tuple(out)

    def leaf_momentum(self, ax, ay, az, leaf_mom=None, device_ptrs=False):
        """sum_i m_i a_i per listed leaf (FP64) of float accelerations."""
        if leaf_mom is None:
            leaf_mom = np.zeros((max(self.stats().listed_leaves, 1), 3))
        fl = HSR_DEVICE_PTRS if device_ptrs else HSR_HOST_PTRS
        check(self.L.hsr_leaf_momentum(self.h, _ptr(ax), _ptr(ay), _ptr(az), _ptr(leaf_mom), fl),
              "hsr_leaf_momentum")
        return leaf_mom

    # -- sub-cycle
    def drift(self, x, y, z, vx, vy, vz, pp, tau, adot, alpha, device_ptrs=False):
        """Particles::map1 in place on x, y, z (float32 arrays or tensors)."""
        n = int(x.shape[0])
        fl = HSR_DEVICE_PTRS if device_ptrs else HSR_HOST_PTRS
        check(self.L.hsr_drift(self.h, n, _ptr(x), _ptr(y), _ptr(z), _ptr(vx), _ptr(vy),
                               _ptr(vz), C.c_float(pp), C.c_float(tau), C.c_float(adot),
                               C.c_float(alpha), fl), "hsr_drift")

    def resort(self, ng, x, y, z, vx=None, vy=None, vz=None, mass=None, phi=None,
               ids=None, mask=None, device_ptrs=False) -> int:
        """Particles::resortParticles in place; returns m_Np_last."""
        n = int(x.shape[0])
        g = np.ascontiguousarray(ng, np.int32)
        inside = C.c_int64(0)
        fl = HSR_DEVICE_PTRS if device_ptrs else HSR_HOST_PTRS
        check(self.L.hsr_resort(self.h, n, g.ctypes.data, _ptr(x), _ptr(y), _ptr(z),
                                _ptr(vx), _ptr(vy), _ptr(vz), _ptr(mass), _ptr(phi),
                                _ptr(ids), _ptr(mask), C.byref(inside), fl), "hsr_resort")
        return int(inside.value)

    def resort_noalloc(self, ng, x, y, z, vx=None, vy=None, vz=None, mass=None, phi=None,
                       ids=None, mask=None) -> int:
        """hsr_resort_noalloc: resortParticles on device arrays over the
        library's static scratch (no device allocation); returns m_Np_last."""
        n = int(x.shape[0])
        g = np.ascontiguousarray(ng, np.int32)
        inside = C.c_int64(0)
        check(self.L.hsr_resort_noalloc(self.h, n, g.ctypes.data, _ptr(x), _ptr(y), _ptr(z),
                                        _ptr(vx), _ptr(vy), _ptr(vz), _ptr(mass), _ptr(phi),
                                        _ptr(ids), _ptr(mask), C.byref(inside),
                                        HSR_DEVICE_PTRS), "hsr_resort_noalloc")
        return int(inside.value)

    def map2(self, params: "SubcycleParams", step_fraction, x, y, z, vx, vy, vz, mass,
             phi=None, ids=None, mask=None, device_ptrs=False) -> int:
        """Particles::map2 (resort, unit masses, tree, lists, kick); returns
        m_Np_last."""
        n = int(x.shape[0])
        inside = C.c_int64(0)
        fl = HSR_DEVICE_PTRS if device_ptrs else HSR_HOST_PTRS
        check(self.L.hsr_map2(self.h, C.byref(params), float(step_fraction), n, _ptr(x),
                              _ptr(y), _ptr(z), _ptr(vx), _ptr(vy), _ptr(vz), _ptr(mass),
                              _ptr(phi), _ptr(ids), _ptr(mask), C.byref(inside), fl),
              "hsr_map2")
        self.n = int(inside.value)
        return int(inside.value)

    def subcycle(self, params: "SubcycleParams", x, y, z, vx, vy, vz, mass, phi=None,
                 ids=None, mask=None, device_ptrs=False) -> int:
        """Particles::subCycle with map2's RCB monopole kick; returns the
        particle count of the last tree (m_Np_last, -1 without kicks)."""
        n = int(x.shape[0])
        inside = C.c_int64(0)
        fl = HSR_DEVICE_PTRS if device_ptrs else HSR_HOST_PTRS
        check(self.L.hsr_subcycle(self.h, C.byref(params), n, _ptr(x), _ptr(y), _ptr(z),
                                  _ptr(vx), _ptr(vy), _ptr(vz), _ptr(mass), _ptr(phi),
                                  _ptr(ids), _ptr(mask), C.byref(inside), fl), "hsr_subcycle")
        self.n = int(inside.value) if inside.value >= 0 else self.n
        return int(inside.value)

    def refresh(self, dom: "Domain", anow, count, x, y, z, vx, vy, vz, mass, phi, ids, mask,
                device_ptrs=False):
        """The one-rank overload refresh (MC3Extras::refreshParticles after
        copyAliveIntoVectors) in place on arrays of capacity len(x) holding
        `count` particles; returns (refreshed count, m_Np_last)."""
        arrs = (x, y, z, vx, vy, vz, mass, phi, ids, mask)
        cap = int(x.shape[0])
        if any(a is None or int(a.shape[0]) != cap for a in arrs):
            raise HsrError("refresh: all ten arrays are required, with one length (the capacity)")
        n = C.c_int64(int(count))
        inside = C.c_int64(0)
        fl = HSR_DEVICE_PTRS if device_ptrs else HSR_HOST_PTRS
        st = self.L.hsr_refresh(self.h, C.byref(dom), C.c_float(anow), C.byref(n), cap,
                                *(_ptr(a) for a in arrs), C.byref(inside), fl)
        if st == HSR_ERR_ARG and n.value > cap:
            raise CapacityError(f"hsr_refresh: {self.L.hsr_last_error().decode()}", int(n.value))
        check(st, "hsr_refresh")
        return int(n.value), int(inside.value)

    # -- PM particle-grid steps
    def cic(self, x, y, z, ng, gpscal=1.0, rho=None, device_ptrs=False):
        """Particles::cic: rho of ng0*ng1*ng2 + 1 floats (the last slot
        collects corners outside the grid)."""
        g = np.ascontiguousarray(ng, np.int32)
        if rho is None:
            rho = np.zeros(int(np.prod(g.astype(np.int64))) + 1, np.float32)
        fl = HSR_DEVICE_PTRS if device_ptrs else HSR_HOST_PTRS
        check(self.L.hsr_cic(self.h, int(x.shape[0]), _ptr(x), _ptr(y), _ptr(z), g.ctypes.data,
                             C.c_float(gpscal), _ptr(rho), fl), "hsr_cic")
        return rho

    def inverse_cic(self, x, y, z, ng, grid, tau, fscal, v, device_ptrs=False):
        """Particles::inverse_cic for one component, v updated in place."""
        g = np.ascontiguousarray(ng, np.int32)
        fl = HSR_DEVICE_PTRS if device_ptrs else HSR_HOST_PTRS
        check(self.L.hsr_inverse_cic(self.h, int(x.shape[0]), _ptr(x), _ptr(y), _ptr(z),
                                     g.ctypes.data, _ptr(grid), C.c_float(tau),
                                     C.c_float(fscal), _ptr(v), fl), "hsr_inverse_cic")
        return v

    # -- introspection
    def stats(self) -> Stats:
        s = Stats()
        check(self.L.hsr_get_stats(self.h, C.byref(s)), "hsr_get_stats")
        return s

    def nodes(self) -> np.ndarray:
        st = self.stats()
        out = np.zeros(max(st.nodes, 1), NODE_DTYPE)
        check(self.L.hsr_get_nodes(self.h, out.ctypes.data, out.shape[0]), "hsr_get_nodes")
        return out[: st.nodes]

    def pseudo_masses(self) -> np.ndarray:
        """(nodes, TDPTS) pseudo-particle masses in canonical node order."""
        st = self.stats()
        T = int(self.params.tdpts)
        out = np.zeros((max(st.nodes, 1), T), np.float32)
        check(self.L.hsr_get_pseudo_masses(self.h, out.ctypes.data, out.size),
              "hsr_get_pseudo_masses")
        return out[: st.nodes]

    def perm(self) -> np.ndarray:
        out = np.zeros(max(self.n, 1), np.int64)
        check(self.L.hsr_get_perm(self.h, out.ctypes.data, out.shape[0]), "hsr_get_perm")
        return out[: self.n]

    def get_lists(self):
        st = self.stats()
        nl, ns = st.listed_leaves, st.segments
        leaves = np.zeros(max(nl, 1), np.int64)
        seg_ptr = np.zeros(nl + 1, np.int64)
        sa = np.zeros(max(ns, 1), np.int64)
        sb = np.zeros(max(ns, 1), np.int64)
        ent = np.zeros(max(nl, 1), np.int64)
        check(self.L.hsr_get_lists(self.h, leaves.ctypes.data, seg_ptr.ctypes.data,
                                   sa.ctypes.data, sb.ctypes.data, ent.ctypes.data,
                                   max(nl, 1), max(ns, 1)), "hsr_get_lists")
        return leaves[:nl], seg_ptr, sa[:ns], sb[:ns], ent[:nl]

    def stream(self) -> int:
        s = C.c_void_p()
        check(self.L.hsr_get_stream(self.h, C.byref(s)), "hsr_get_stream")
        return s.value or 0

    def set_tuning(self, **kw):
        t = Tuning()
        check(self.L.hsr_get_tuning(self.h, C.byref(t)), "hsr_get_tuning")
        for k, v in kw.items():
            setattr(t, k, int(v))
        check(self.L.hsr_set_tuning(self.h, C.byref(t)), "hsr_set_tuning")


def domain(rL, oL, ng) -> Domain:
    """hsr_domain_init: Domain::initialize for one rank (no device needed)."""
    d = Domain(rL=float(rL), oL_in=float(oL), ng=int(ng))
    check(load().hsr_domain_init(C.byref(d)), "hsr_domain_init")
    return d
