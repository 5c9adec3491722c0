"""Write golden vectors from the reference's own compiled sources.

Runs only where the reference tree is mounted (it builds oracle/_ref from
ForceLaw.cxx and BGQCM.c where they lie).  Output:
  ref_cm.npz        inputs and outputs of the reference's generic cm()
                    (BGQCM.c:181-212) on ranges chosen to make the double
                    sums round (mixed magnitudes, ties, long runs)
  ref_forcelaw.npz  ForceLawSR(FGridEvalPoly) and ForceLawSR(FGridEvalFit)
                    f_over_r, FGridEvalFit::eval, FGridEvalPoly::eval and
                    FGridEvalInterp(1024)::eval on an r^2 grid, plus
                    FGrid::rmax (ForceLaw.cxx); ForceLawSR(FGridEvalInterp(n))
                    f_over_r for n in INTERP_SIZES on the same grid and on
                    each table's nodes and midpoints, and FGrid's r^2
                    tables of those sizes
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402


def cm_cases(rng):
    cases = []
    # box coordinates with unit masses (the production case)
    for n in (1, 2, 7, 64, 65, 1000, 5000):
        p = (rng.random((n, 3)) * 262.4).astype(np.float32)
        cases.append((p, np.ones(n, np.float32)))
    # mixed magnitudes: tiny and large coordinates force double rounding
    for n in (300, 3000):
        e = rng.integers(-40, 10, size=(n, 3))
        p = (rng.random((n, 3)) * np.exp2(e)).astype(np.float32)
        cases.append((p, np.ones(n, np.float32)))
    # non-unit masses and negative coordinates
    for n in (500, 2500):
        p = ((rng.random((n, 3)) - 0.3) * 100).astype(np.float32)
        w = (0.5 + rng.random(n)).astype(np.float32)
        cases.append((p, w))
    # exact half-ulp ties: quarter-unit grid values
    n = 4096
    p = (rng.integers(0, 4096, size=(n, 3)) * 0.25 + 2.0**24).astype(np.float32)
    cases.append((p, np.ones(n, np.float32)))
    return cases


INTERP_SIZES = (2, 97, 1024, 4096)


def interp_vectors(R, r2, rsm):
    """ForceLawSR(FGridEvalInterp(n)) on the r^2 grid and on n's own table
    nodes and midpoints (where the index rounds), with the tables."""
    out = {}
    for n in INTERP_SIZES:
        tr2 = np.zeros(n, np.float32)
        tf = np.zeros(n, np.float32)
        R.hsrref_interp_table(n, tr2, tf)
        mid = ((tr2[:-1].astype(np.float64) + tr2[1:]) / 2).astype(np.float32)
        pts = np.concatenate([r2, tr2, mid, np.nextafter(tr2, np.float32(np.inf))]).astype(np.float32)
        out[f"interp{n}_r2"] = pts
        out[f"sr_interp{n}"] = np.array([R.hsrref_sr_interp_f_over_r(float(v), float(rsm), n)
                                         for v in pts], np.float32)
        out[f"table{n}_r2"] = tr2
        out[f"table{n}_f"] = tf
    return out


def main():
    R = O.ref()
    if R is None:
        O.build()
        R = O.ref()
    if R is None:
        sys.exit("oracle/_ref/libhsrref.so is not available (needs the reference tree)")
    rng = np.random.default_rng(20261202)
    xs, ws, offs, outs = [], [], [0], []
    for p, w in cm_cases(rng):
        x, y, z = (np.ascontiguousarray(p[:, k]) for k in range(3))
        xmin = np.zeros(3, np.float32); xmax = np.zeros(3, np.float32); xc = np.zeros(3, np.float32)
        R.hsrref_cm(p.shape[0], x, y, z, w, xmin, xmax, xc)
        xs.append(p); ws.append(w); offs.append(offs[-1] + p.shape[0])
        outs.append(np.concatenate([xmin, xmax, xc]))
    np.savez_compressed(os.path.join(HERE, "ref_cm.npz"), pos=np.concatenate(xs),
                        mass=np.concatenate(ws), offsets=np.array(offs, np.int64),
                        out=np.stack(outs))
    r2 = np.concatenate([np.array([0.0, 1e-6, 1e-3, 0.01], np.float32),
                         np.linspace(0.0, 12.0, 1201, dtype=np.float32),
                         np.float32(3.116326355) ** 2 + np.array([-1e-6, 0, 1e-6], np.float32)]).astype(np.float32)
    rsm = np.float32(0.1)
    f = lambda fn, *a: np.array([fn(float(v), *a) for v in r2], np.float32)
    np.savez_compressed(os.path.join(HERE, "ref_forcelaw.npz"), r2=r2, rsm=rsm,
                        sr_poly6=f(R.hsrref_sr_poly_f_over_r, float(rsm)),
                        sr_fit=f(R.hsrref_sr_fit_f_over_r, float(rsm)),
                        fgor=f(R.hsrref_fgor_r2), poly6=f(R.hsrref_poly_eval),
                        interp1024=f(R.hsrref_interp_eval),
                        rmax=np.float32(R.hsrref_fgrid_rmax()),
                        **interp_vectors(R, r2, rsm))
    print("wrote", os.listdir(HERE))


if __name__ == "__main__":
    main()
