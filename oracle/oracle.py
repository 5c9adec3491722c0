"""ctypes front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Only tests/,
the smoke entry point and bench.py's cpu_baseline leg may import this.

Each wrapper follows the C function of the same name in hsr_oracle.c, which
in turn cites the reference file:line it restates.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libhsrref.so")

LAW_STEP16, LAW_SR_POLY5, LAW_SR_POLY6, LAW_NEWTON, LAW_SR_FIT, LAW_SR_INTERP = 0, 1, 2, 3, 4, 5


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


class Domain(C.Structure):
    _fields_ = [("rL", C.c_float), ("oL_in", C.c_float), ("ng", C.c_int32),
                ("ng_overload", C.c_int32), ("ng_local_total", C.c_int32),
                ("oL", C.c_float), ("phys2grid_pos", C.c_float),
                ("grid2phys_pos", C.c_float), ("phys2grid_vel", C.c_float),
                ("grid2phys_vel", C.c_float)]


class Params(C.Structure):
    _fields_ = [("nd", C.c_int64), ("fsrrmax", C.c_float), ("rsm", C.c_float),
                ("tan_oa", C.c_float), ("ppc", C.c_float),
                ("min_loc", C.c_float * 3), ("max_loc", C.c_float * 3),
                ("min_force", C.c_float * 3), ("max_force", C.c_float * 3),
                ("tdpts", C.c_int32), ("n_interp", C.c_int32), ("ppm_t", C.c_void_p)]


_lib = None
_ref = None


def build():
    """Compile liboracle.so (and _ref/libhsrref.so when the reference tree
    is present)."""
    subprocess.check_call(["make", "-s", "-C", HERE, "all"])
    if os.path.isdir("/root/reference/src/halo_finder"):
        subprocess.check_call(["make", "-s", "-C", HERE, "ref"])


def set_threads(n: int) -> int:
    """OpenMP threads of the oracle's parallel loops; returns the previous."""
    return int(lib().hsro_set_threads(int(n)))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = load(LIB_PATH)
    return _lib


def use(library):
    """Make the wrappers below call `library` (a handle from load()) and
    return the one they called before; the mutation tests swap in oracles
    with planted mistakes this way."""
    global _lib
    old = _lib
    _lib = library
    return old


def load(path):
    """Load an oracle library built from hsr_oracle.c and declare its
    functions."""
    L = C.CDLL(path)
    P = np.ctypeslib.ndpointer
    f32 = P(np.float32, flags="C")
    f64 = P(np.float64, flags="C")
    i64 = P(np.int64, flags="C")
    u16 = P(np.uint16, flags="C")
    nodes = P(NODE_DTYPE, flags="C")
    L.hsro_set_threads.argtypes = [C.c_int]
    L.hsro_set_threads.restype = C.c_int
    L.hsro_cm.argtypes = [C.c_int64, f32, f32, f32, f32, f32, f32, f32]
    L.hsro_partition.argtypes = [C.c_int64] + [f32] * 8 + [i64, u16, C.c_float]
    L.hsro_partition.restype = C.c_int64
    L.hsro_build.argtypes = [C.POINTER(Params), C.c_int64] + [f32] * 8 + [i64, u16, nodes, C.c_int64]
    L.hsro_build.restype = C.c_int64
    L.hsro_leaf_order.argtypes = [C.POINTER(Params), nodes, C.c_int64, i64, C.c_int64]
    L.hsro_leaf_order.restype = C.c_int64
    L.hsro_list.argtypes = [C.POINTER(Params), nodes, C.c_int64, C.c_int64, i64, i64, C.c_int64, C.POINTER(C.c_int64)]
    L.hsro_list.restype = C.c_int64
    L.hsro_lists_all.argtypes = [C.POINTER(Params), nodes, C.c_int64, i64, C.c_int64, i64, i64, i64, C.c_int64, i64]
    L.hsro_lists_all.restype = C.c_int64
    L.hsro_force_f32.argtypes = [C.c_int, C.POINTER(Params), nodes, f32, f32, f32, f32, i64, C.c_int64, i64, i64, i64, f32, f32, f32]
    L.hsro_force_step16.argtypes = [C.POINTER(Params), nodes, f32, f32, f32, f32, i64, C.c_int64, i64, i64, i64, f32, f32, f32]
    L.hsro_force_f64.argtypes = [C.POINTER(Params), nodes, f32, f32, f32, f32, i64, C.c_int64, i64, i64, i64, f64, f64, f64, C.c_void_p]
    L.hsro_law_f32.argtypes = [C.c_int, C.c_float, C.c_float, C.c_float]
    L.hsro_law_f32.restype = C.c_float
    L.hsro_direct_f64.argtypes = [C.c_int64, f32, f32, f32, f32, i64, C.c_int64, C.c_float, C.c_float, f64, f64, f64]
    L.hsro_adjoint_f64.argtypes = [C.POINTER(Params), nodes, f64, f64, f64, f64, i64, C.c_int64, i64, i64, i64, f64, f64, f64, f64, f64, f64]
    L.hsro_objective_f64.argtypes = [C.POINTER(Params), nodes, f64, f64, f64, f64, i64, C.c_int64, i64, i64, i64, f64, f64, f64]
    L.hsro_objective_f64.restype = C.c_double
    L.hsro_force_f64d.argtypes = [C.POINTER(Params), nodes, f64, f64, f64, f64, i64, C.c_int64, i64, i64, i64, f64, f64, f64]
    L.hsro_cell_sort_perm.argtypes = [C.c_int64, f32, f32, f32, P(np.int32, flags="C"), i64]
    L.hsro_cell_sort_perm.restype = C.c_int64
    L.hsro_pppts.argtypes = [C.c_int, C.c_float, f32, f32, f32, f32]
    i32 = P(np.int32, flags="C")
    L.hsro_cic.argtypes = [C.c_int64, f32, f32, f32, i32, C.c_float, f32]
    L.hsro_inverse_cic.argtypes = [C.c_int64, f32, f32, f32, i32, f32, C.c_float,
                                   C.c_float, f32]
    L.hsro_map1_factor.argtypes = [C.c_float] * 4
    L.hsro_map1_factor.restype = C.c_float
    L.hsro_drift.argtypes = [C.c_int64] + [f32] * 6 + [C.c_float]
    L.hsro_map2_coeff.argtypes = [C.c_float, C.c_double, C.c_double, C.c_double]
    L.hsro_map2_coeff.restype = C.c_float
    L.hsro_domain_init.argtypes = [C.POINTER(Domain)]
    L.hsro_refresh.argtypes = [C.POINTER(Domain), C.c_float, C.c_int64, C.c_int64] + \
        [f32] * 8 + [i64, u16, C.POINTER(C.c_int64)]
    L.hsro_refresh.restype = C.c_int64
    L.hsro_law_f32_vec.argtypes = [C.c_int, C.c_int64, f32, f32, C.c_float, C.c_float, C.c_int]
    L.hsro_fgrid_table.argtypes = [C.c_int, f32, f32]
    L.hsro_law_coeffs.argtypes = [C.c_int, P(np.float32, flags="C")]
    L.hsro_law_coeffs.restype = C.c_int
    return L


def ref():
    """The reference's own ForceLaw.cxx + BGQCM.c (oracle/_ref), or None
    when it was not built (it is only buildable where the reference tree
    is mounted)."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF_PATH):
            return None
        R = C.CDLL(REF_PATH)
        P = np.ctypeslib.ndpointer
        f32 = P(np.float32, flags="C")
        R.hsrref_cm.argtypes = [C.c_int64] + [f32] * 7
        for name in ("hsrref_sr_poly_f_over_r", "hsrref_sr_fit_f_over_r"):
            getattr(R, name).argtypes = [C.c_float, C.c_float]
            getattr(R, name).restype = C.c_float
        for name in ("hsrref_fgor_r2", "hsrref_poly_eval", "hsrref_newton_f_over_r",
                     "hsrref_interp_eval"):
            getattr(R, name).argtypes = [C.c_float]
            getattr(R, name).restype = C.c_float
        R.hsrref_fgrid_rmax.restype = C.c_float
        R.hsrref_sr_interp_f_over_r.argtypes = [C.c_float, C.c_float, C.c_int]
        R.hsrref_sr_interp_f_over_r.restype = C.c_float
        R.hsrref_interp_table.argtypes = [C.c_int, f32, f32]
        _ref = R
    return _ref


@dataclasses.dataclass
class TreeParams:
    """Constructor arguments of RCBForceTree (RCBForceTree.h:96-124) that
    shape the tree, the lists and the kernel."""
    min_loc: np.ndarray
    max_loc: np.ndarray
    min_force: np.ndarray
    max_force: np.ndarray
    fsrrmax: float = 3.12
    rsm: float = 0.1
    oa: float = 0.1
    nd: int = 128
    ppc: float = 0.9
    tdpts: int = 1          # 1: RCBMonopoleForceTree, 12: RCBQuadrupoleForceTree
    n_interp: int = 0       # LAW_SR_INTERP: FGridEvalInterp's table size
    tan: float = None       # tanOpeningAngle itself, instead of tanf(oa)

    def tan_oa(self) -> float:
        # tanf(oa) in float, as RCBForceTree.cxx:381 evaluates it
        if self.tan is not None:
            return float(np.float32(self.tan))
        return tanf(self.oa)

    def ctypes(self, tree=None) -> Params:
        """The C parameter block; a tree with T > 1 lends its pseudo masses."""
        p = Params()
        p.tdpts = int(self.tdpts)
        p.n_interp = int(self.n_interp)
        if tree is not None and getattr(tree, "ppm_t", None) is not None:
            p.ppm_t = tree.ppm_t.ctypes.data
        p.nd = int(self.nd)
        p.fsrrmax = float(np.float32(self.fsrrmax))
        p.rsm = float(np.float32(self.rsm))
        p.tan_oa = self.tan_oa()
        p.ppc = float(np.float32(self.ppc))
        for k in range(3):
            p.min_loc[k] = float(self.min_loc[k])
            p.max_loc[k] = float(self.max_loc[k])
            p.min_force[k] = float(self.min_force[k])
            p.max_force[k] = float(self.max_force[k])
        return p


def tanf(x: float) -> float:
    """tanf as glibc evaluates it (the reference calls tanf(oa) once,
    RCBForceTree.cxx:381)."""
    libm = C.CDLL("libm.so.6")
    libm.tanf.argtypes = [C.c_float]
    libm.tanf.restype = C.c_float
    return float(libm.tanf(C.c_float(x)))


@dataclasses.dataclass
class Tree:
    nodes: np.ndarray      # NODE_DTYPE, canonical (unthreaded allocation) ids
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    vx: np.ndarray
    vy: np.ndarray
    vz: np.ndarray
    mass: np.ndarray
    phi: np.ndarray
    id: np.ndarray
    mask: np.ndarray
    ppm_t: np.ndarray = None   # T > 1: (nodes, T) pseudo masses

    def ppm(self, T=1) -> np.ndarray:
        """Pseudo masses as (nodes, T)."""
        return self.nodes["ppm"].reshape(-1, 1) if self.ppm_t is None else self.ppm_t


def build_tree(tp: TreeParams, x, y, z, mass=None, vx=None, vy=None, vz=None,
               phi=None, ids=None, mask=None) -> Tree:
    """RCBForceTree constructor's build phase on copies of the inputs."""
    n = x.shape[0]
    f = lambda a, d=0.0: (np.full(n, d, np.float32) if a is None
                          else np.ascontiguousarray(a, np.float32).copy())
    X, Y, Z = f(x), f(y), f(z)
    M = f(mass, 1.0)
    VX, VY, VZ, PHI = f(vx), f(vy), f(vz), f(phi)
    ID = np.arange(n, dtype=np.int64) if ids is None else np.ascontiguousarray(ids, np.int64).copy()
    MASK = np.zeros(n, np.uint16) if mask is None else np.ascontiguousarray(mask, np.uint16).copy()
    cap = max(16, 4 * (n // max(1, tp.nd // 2) + 2) + 64)
    while True:
        nodes = np.zeros(cap, NODE_DTYPE)
        ppm_t = np.zeros((cap, tp.tdpts), np.float32) if tp.tdpts > 1 else None
        p = tp.ctypes()
        if ppm_t is not None:
            p.ppm_t = ppm_t.ctypes.data
        # rebuild from pristine inputs each attempt
        X2, Y2, Z2, M2 = X.copy(), Y.copy(), Z.copy(), M.copy()
        VX2, VY2, VZ2, PHI2, ID2, MASK2 = VX.copy(), VY.copy(), VZ.copy(), PHI.copy(), ID.copy(), MASK.copy()
        nn = lib().hsro_build(C.byref(p), n, X2, Y2, Z2, VX2, VY2, VZ2, M2,
                              PHI2, ID2, MASK2, nodes, cap)
        assert nn != -2, "unsupported tdpts"
        if nn >= 0:
            return Tree(nodes[:nn].copy(), X2, Y2, Z2, VX2, VY2, VZ2, M2,
                        PHI2, ID2, MASK2, None if ppm_t is None else ppm_t[:nn].copy())
        cap *= 2


def leaf_order(tp: TreeParams, tree: Tree) -> np.ndarray:
    nn = tree.nodes.shape[0]
    out = np.zeros(nn + 1, np.int64)
    p = tp.ctypes()
    k = lib().hsro_leaf_order(C.byref(p), tree.nodes, nn, out, out.shape[0])
    assert k >= 0
    return out[:k].copy()


@dataclasses.dataclass
class Lists:
    leaves: np.ndarray     # target leaf node ids, in leaf order
    seg_ptr: np.ndarray    # CSR over segments (len nleaves+1)
    seg_a: np.ndarray      # particle offset, or node id for a pseudo-particle
    seg_b: np.ndarray      # particle count, or -1 for a pseudo-particle
    entries: np.ndarray    # list length per leaf

    def pairs(self, tree: Tree) -> int:
        return int(np.sum(tree.nodes["count"][self.leaves] * self.entries))


def build_lists(tp: TreeParams, tree: Tree, leaves=None) -> Lists:
    if leaves is None:
        leaves = leaf_order(tp, tree)
    leaves = np.ascontiguousarray(leaves, np.int64)
    nl = leaves.shape[0]
    seg_ptr = np.zeros(nl + 1, np.int64)
    ent = np.zeros(max(nl, 1), np.int64)
    p = tp.ctypes()
    cap = max(64, nl * 64)
    while True:
        sa = np.zeros(cap, np.int64)
        sb = np.zeros(cap, np.int64)
        tot = lib().hsro_lists_all(C.byref(p), tree.nodes, tree.nodes.shape[0],
                                   leaves, nl, seg_ptr, sa, sb, cap, ent)
        if tot >= 0:
            return Lists(leaves, seg_ptr.copy(), sa[:tot].copy(), sb[:tot].copy(), ent[:nl].copy())
        cap = int(seg_ptr[nl]) + 16


def forces_f32(tp: TreeParams, tree: Tree, lists: Lists, law: int = LAW_STEP16):
    n = tree.x.shape[0]
    ax = np.zeros(n, np.float32); ay = np.zeros(n, np.float32); az = np.zeros(n, np.float32)
    p = tp.ctypes(tree)
    lib().hsro_force_f32(law, C.byref(p), tree.nodes, tree.x, tree.y, tree.z,
                         tree.mass, lists.leaves, lists.leaves.shape[0],
                         lists.seg_ptr, lists.seg_a, lists.seg_b, ax, ay, az)
    return ax, ay, az


def forces_step16(tp: TreeParams, tree: Tree, lists: Lists):
    n = tree.x.shape[0]
    ax = np.zeros(n, np.float32); ay = np.zeros(n, np.float32); az = np.zeros(n, np.float32)
    p = tp.ctypes(tree)
    lib().hsro_force_step16(C.byref(p), tree.nodes, tree.x, tree.y, tree.z,
                            tree.mass, lists.leaves, lists.leaves.shape[0],
                            lists.seg_ptr, lists.seg_a, lists.seg_b, ax, ay, az)
    return ax, ay, az


def forces_f64(tp: TreeParams, tree: Tree, lists: Lists, with_momentum=False):
    n = tree.x.shape[0]
    ax = np.zeros(n); ay = np.zeros(n); az = np.zeros(n)
    mom = np.zeros((lists.leaves.shape[0], 3)) if with_momentum else None
    p = tp.ctypes(tree)
    lib().hsro_force_f64(C.byref(p), tree.nodes, tree.x, tree.y, tree.z,
                         tree.mass, lists.leaves, lists.leaves.shape[0],
                         lists.seg_ptr, lists.seg_a, lists.seg_b, ax, ay, az,
                         mom.ctypes.data if with_momentum else None)
    return (ax, ay, az, mom) if with_momentum else (ax, ay, az)


def direct_f64(x, y, z, m, targets, rsm=0.1, fsrrmax=3.12):
    targets = np.ascontiguousarray(targets, np.int64)
    k = targets.shape[0]
    ax = np.zeros(k); ay = np.zeros(k); az = np.zeros(k)
    lib().hsro_direct_f64(x.shape[0], np.ascontiguousarray(x, np.float32),
                          np.ascontiguousarray(y, np.float32),
                          np.ascontiguousarray(z, np.float32),
                          np.ascontiguousarray(m, np.float32), targets, k,
                          float(np.float32(rsm)), float(np.float32(fsrrmax)),
                          ax, ay, az)
    return ax, ay, az


def law_f32(law: int, r2, rsm=0.1, fsrrmax=3.12, n_interp=0):
    r2 = np.ascontiguousarray(r2, np.float32).ravel()
    out = np.zeros_like(r2)
    lib().hsro_law_f32_vec(law, r2.shape[0], r2, out, float(np.float32(rsm)),
                           float(np.float32(fsrrmax)), int(n_interp))
    return out


def law_coeffs(law: int) -> np.ndarray:
    """The polynomial coefficients a0.. the oracle uses for a law."""
    out = np.zeros(7, np.float32)
    k = lib().hsro_law_coeffs(law, out)
    return out[:k].copy()


def fgrid_table(n: int):
    """FGrid::fgor_r2_interp's table (r2, f) of n entries."""
    r2 = np.zeros(n, np.float32)
    f = np.zeros(n, np.float32)
    lib().hsro_fgrid_table(n, r2, f)
    return r2, f


def cm(x, y, z, m):
    xmin = np.zeros(3, np.float32); xmax = np.zeros(3, np.float32); xc = np.zeros(3, np.float32)
    lib().hsro_cm(x.shape[0], np.ascontiguousarray(x, np.float32),
                  np.ascontiguousarray(y, np.float32),
                  np.ascontiguousarray(z, np.float32),
                  np.ascontiguousarray(m, np.float32), xmin, xmax, xc)
    return xmin, xmax, xc


def partition(key, pv, *others):
    """hsro_partition on copies; returns (is, key', others'...)."""
    n = key.shape[0]
    arrs = [np.ascontiguousarray(key, np.float32).copy()]
    fl = [np.ascontiguousarray(o, np.float32).copy() for o in others[:7]]
    while len(fl) < 7:
        fl.append(np.zeros(n, np.float32))
    ids = np.arange(n, dtype=np.int64)
    mask = np.zeros(n, np.uint16)
    is_ = lib().hsro_partition(n, arrs[0], *fl, ids, mask, float(np.float32(pv)))
    return is_, arrs[0], ids


def cell_sort_perm(x, y, z, ng):
    perm = np.zeros(x.shape[0], np.int64)
    inside = lib().hsro_cell_sort_perm(x.shape[0], np.ascontiguousarray(x, np.float32),
                                       np.ascontiguousarray(y, np.float32),
                                       np.ascontiguousarray(z, np.float32),
                                       np.asarray(ng, np.int32), perm)
    return perm, inside


def pppts(tdpts, tdr, xc):
    """(T, 3) positions of a node's pseudo-particles."""
    out = [np.zeros(tdpts, np.float32) for _ in range(3)]
    lib().hsro_pppts(tdpts, C.c_float(tdr), np.ascontiguousarray(xc, np.float32), *out)
    return np.stack(out, 1)


def cic(x, y, z, ng, gpscal=1.0):
    """Particles::cic: rho of ng0*ng1*ng2 + 1 cells (last = outside)."""
    g = np.asarray(ng, np.int32)
    rho = np.zeros(int(np.prod(g.astype(np.int64))) + 1, np.float32)
    f = lambda a: np.ascontiguousarray(a, np.float32)
    lib().hsro_cic(x.shape[0], f(x), f(y), f(z), g, C.c_float(gpscal), rho)
    return rho


def inverse_cic(x, y, z, ng, grid, tau, fscal, v):
    """Particles::inverse_cic for one component; returns v + the kick."""
    g = np.asarray(ng, np.int32)
    out = np.ascontiguousarray(v, np.float32).copy()
    f = lambda a: np.ascontiguousarray(a, np.float32)
    lib().hsro_inverse_cic(x.shape[0], f(x), f(y), f(z), g, f(grid), C.c_float(tau),
                           C.c_float(fscal), out)
    return out


def map1_factor(pp, tau, adot, alpha) -> float:
    return float(lib().hsro_map1_factor(C.c_float(pp), C.c_float(tau), C.c_float(adot),
                                        C.c_float(alpha)))


def drift(x, y, z, vx, vy, vz, factor):
    """Particles::map1's update on copies; returns (x, y, z)."""
    X, Y, Z = (np.ascontiguousarray(a, np.float32).copy() for a in (x, y, z))
    V = [np.ascontiguousarray(a, np.float32) for a in (vx, vy, vz)]
    lib().hsro_drift(X.shape[0], X, Y, Z, V[0], V[1], V[2], C.c_float(factor))
    return X, Y, Z


def map2_coeff(gpscal, fscal, tau, step_fraction) -> float:
    return float(lib().hsro_map2_coeff(C.c_float(gpscal), fscal, tau, step_fraction))


POSVEL = ("x", "y", "z", "vx", "vy", "vz", "mass", "phi")


def resort(state: dict, ng):
    """Particles::resortParticles on a dict of the ten arrays (copies);
    returns (new state, m_Np_last)."""
    perm, inside = cell_sort_perm(state["x"], state["y"], state["z"], ng)
    out = {}
    for k, a in state.items():
        b = np.empty_like(a)
        b[perm] = a
        out[k] = b
    return out, int(inside)


def domain(rL, oL, ng) -> Domain:
    """Domain::initialize for one rank (Domain.cxx:34-84)."""
    d = Domain(rL=float(rL), oL_in=float(oL), ng=int(ng))
    lib().hsro_domain_init(C.byref(d))
    return d


def refresh(state: dict, count: int, rL, oL, ng, anow, sort=True):
    """MC3Extras' particle refresh on one rank (copyAliveIntoVectors,
    ParticleExchange, copyAllFromVectors, then resortParticles unless sort
    is False) on copies of the ten arrays, whose length is the capacity.
    Returns (state, count, m_Np_last, alive); m_Np_last is -1 unsorted."""
    d = domain(rL, oL, ng)
    st = {a: np.ascontiguousarray(b).copy() for a, b in state.items()}
    cap = st["x"].shape[0]
    alive = C.c_int64(0)
    n = lib().hsro_refresh(C.byref(d), C.c_float(anow), int(count), cap,
                           *(st[a] for a in POSVEL), st["id"], st["mask"], C.byref(alive))
    if n < 0:
        raise ValueError("refreshed particles exceed the capacity")
    if not sort:
        return st, int(n), -1, int(alive.value)
    sub = {a: b[:n].copy() for a, b in st.items()}
    sub, inside = resort(sub, [d.ng_local_total] * 3)
    for a in st:
        st[a][:n] = sub[a]
    return st, int(n), inside, int(alive.value)


def substep(state: dict, sp: dict, tp_kw: dict):
    """One sub-step of Particles::subCycle with map2's RCB monopole kick
    (Particles.cxx:1188-1198, 1205-1280) on copies of the ten arrays:
    drift, resort, masses set to 1, tree over the particles inside the
    grid, lists and Step16 kick of the force box, drift.  sp holds nsub,
    ng, alpha, gpscal, edge, pp, tau, tau2, adot, fscal; tp_kw the tree's
    fsrrmax, rsm, oa, nd.  Returns (state, m_Np_last, tree, lists)."""
    sf = 1.0 / sp["nsub"]
    k = map1_factor(np.float32(sp["pp"]), np.float32(sf * sp["tau2"]),
                    np.float32(sp["adot"]), sp["alpha"])
    c = map2_coeff(sp["gpscal"], sp["fscal"], sp["tau"], sf)
    st = {a: b.copy() for a, b in state.items()}
    st["x"], st["y"], st["z"] = drift(st["x"], st["y"], st["z"], st["vx"], st["vy"], st["vz"], k)
    st, inside = resort(st, sp["ng"])
    st["mass"][:] = 1.0
    ng = sp["ng"]
    gmax = np.float32(max(ng))
    edge = np.float32(sp["edge"])
    tp = TreeParams(np.zeros(3, np.float32), np.full(3, gmax, np.float32),
                    np.full(3, edge, np.float32),
                    np.array([np.float32(1.0 * g - float(edge)) for g in ng], np.float32),
                    **tp_kw)
    sl = slice(0, inside)
    tree = build_tree(tp, st["x"][sl], st["y"][sl], st["z"][sl], st["mass"][sl],
                      st["vx"][sl], st["vy"][sl], st["vz"][sl], st["phi"][sl],
                      st["id"][sl], st["mask"][sl])
    for a in ("x", "y", "z", "vx", "vy", "vz", "mass", "phi", "id", "mask"):
        st[a][sl] = getattr(tree, a)
    lists = build_lists(tp, tree)
    fx, fy, fz = forces_step16(tp, tree, lists)
    cf = np.float32(c)
    for a, f in (("vx", fx), ("vy", fy), ("vz", fz)):
        v = st[a][sl]
        v += cf * f   # zero force outside the listed leaves leaves v unchanged
        st[a][sl] = v
    st["x"], st["y"], st["z"] = drift(st["x"], st["y"], st["z"], st["vx"], st["vy"], st["vz"], k)
    return st, inside, tree, lists


def _d(a):
    return np.ascontiguousarray(a, np.float64)


def adjoint_f64(tp: TreeParams, tree: Tree, lists: Lists, lam, pos=None):
    """FP64 dJ/dx for J = sum lam . F (lam: n x 3, sorted order)."""
    n = tree.x.shape[0]
    x, y, z = (tree.x, tree.y, tree.z) if pos is None else pos
    g = [np.zeros(n) for _ in range(3)]
    p = tp.ctypes()
    lib().hsro_adjoint_f64(C.byref(p), tree.nodes, _d(x), _d(y), _d(z), _d(tree.mass),
                           lists.leaves, lists.leaves.shape[0], lists.seg_ptr,
                           lists.seg_a, lists.seg_b, _d(lam[:, 0]), _d(lam[:, 1]),
                           _d(lam[:, 2]), *g)
    return np.stack(g, 1)


def objective_f64(tp: TreeParams, tree: Tree, lists: Lists, lam, pos):
    """J = sum_i lam_i . F_i in FP64 at positions pos = (x, y, z) (sorted
    order), topology and lists frozen, pseudo-particles at the FP64 centre
    of mass of their members."""
    p = tp.ctypes()
    return lib().hsro_objective_f64(C.byref(p), tree.nodes, _d(pos[0]), _d(pos[1]),
                                    _d(pos[2]), _d(tree.mass), lists.leaves,
                                    lists.leaves.shape[0], lists.seg_ptr, lists.seg_a,
                                    lists.seg_b, _d(lam[:, 0]), _d(lam[:, 1]),
                                    _d(lam[:, 2]))


def force_f64_at(tp: TreeParams, nodes, lists: Lists, pos, mass):
    """FP64 forces at double positions pos = (x, y, z) (sorted order) with
    frozen topology; pseudo-particles at the FP64 centre of mass."""
    n = pos[0].shape[0]
    a = [np.zeros(n) for _ in range(3)]
    p = tp.ctypes()
    lib().hsro_force_f64d(C.byref(p), np.ascontiguousarray(nodes), _d(pos[0]), _d(pos[1]),
                          _d(pos[2]), _d(mass), lists.leaves, lists.leaves.shape[0],
                          lists.seg_ptr, lists.seg_a, lists.seg_b, *a)
    return np.stack(a, 1)


def adjoint_f64_at(tp: TreeParams, nodes, lists: Lists, pos, mass, lam):
    """FP64 dJ/dx (J = sum lam . F) at double positions, frozen topology."""
    n = pos[0].shape[0]
    g = [np.zeros(n) for _ in range(3)]
    p = tp.ctypes()
    lib().hsro_adjoint_f64(C.byref(p), np.ascontiguousarray(nodes), _d(pos[0]), _d(pos[1]),
                           _d(pos[2]), _d(mass), lists.leaves, lists.leaves.shape[0],
                           lists.seg_ptr, lists.seg_a, lists.seg_b, _d(lam[:, 0]),
                           _d(lam[:, 1]), _d(lam[:, 2]), *g)
    return np.stack(g, 1)
