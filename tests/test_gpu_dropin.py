"""GPU test of the C++ drop-in (hacc-coral_amd/host/RCBForceTree.h): the
reference caller pattern -- construct, delete -- permutes the caller's arrays
exactly as the oracle and kicks v by fcoeff * F."""
import os
import subprocess
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "hacc-coral_amd", "bin", "test_dropin")


LAWS = {"step16": (O.LAW_STEP16, 0), "poly5": (O.LAW_SR_POLY5, 0),
        "poly6": (O.LAW_SR_POLY6, 0), "fit": (O.LAW_SR_FIT, 0),
        "interp1024": (O.LAW_SR_INTERP, 1024), "interp97": (O.LAW_SR_INTERP, 97)}


@pytest.mark.parametrize("tdpts,law", [(1, "step16"), (12, "step16"), (1, "poly5"),
                                       (1, "poly6"), (1, "fit"), (1, "interp1024"),
                                       (1, "interp97"), (12, "fit")])
def test_cxx_dropin_constructor(tmp_path, tdpts, law):
    """The Step16 kernel (the BG/Q build) and, through setKernel(ForceLaw),
    nbody1 with ForceLawSR over each grid-force evaluator: FGridEvalPoly of
    order 5 and 6, FGridEvalFit (the reference's default law,
    Particles.cxx:173-174) and FGridEvalInterp."""
    n, box, nd, seed, fcoeff = 30000, 14.0, 64, 7, 0.5
    print(f"case C++ drop-in n={n} nd={nd} TDPTS={tdpts} law={law}", flush=True)
    out = tmp_path / "dropin.bin"
    subprocess.run([EXE, str(n), str(box), str(nd), str(seed), str(out), str(tdpts), law],
                   check=True, timeout=120)
    raw = np.fromfile(out, dtype=np.uint8)
    f = raw[: 9 * 4 * n].view(np.float32).reshape(9, n)
    ids = raw[9 * 4 * n:].view(np.int64)
    x0, y0, z0, x, y, z, vx, vy, vz = f
    law_id, n_interp = LAWS[law]
    tp = O.TreeParams(np.zeros(3, np.float32), np.full(3, box, np.float32),
                      np.full(3, 1.0, np.float32), np.full(3, box - 1.0, np.float32),
                      fsrrmax=float(np.float32(3.116326355)), rsm=0.1, oa=0.1, nd=nd,
                      tdpts=tdpts, n_interp=n_interp)
    tree = O.build_tree(tp, x0, y0, z0)
    assert np.array_equal(ids, tree.id)
    assert np.array_equal(x.view(np.uint32), tree.x.view(np.uint32))
    lists = O.build_lists(tp, tree)
    if law == "step16":
        rx, ry, rz = O.forces_step16(tp, tree, lists)
    else:  # nbody1: v += fcoeff * m_i * sum m_j f_over_r(r2) d, unit masses
        rx, ry, rz = O.forces_f32(tp, tree, lists, law=law_id)
    V = np.stack([vx, vy, vz], 1)
    Rv = fcoeff * np.stack([rx, ry, rz], 1)
    err = np.sqrt(np.sum((V - Rv) ** 2) / np.sum(Rv ** 2))
    assert err <= 1e-5, err
