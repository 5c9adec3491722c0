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


@pytest.mark.parametrize("tdpts", [1, 12])
def test_cxx_dropin_constructor(tmp_path, tdpts):
    n, box, nd, seed, fcoeff = 30000, 14.0, 64, 7, 0.5
    print(f"case C++ drop-in n={n} nd={nd} TDPTS={tdpts}", flush=True)
    out = tmp_path / "dropin.bin"
    subprocess.run([EXE, str(n), str(box), str(nd), str(seed), str(out), str(tdpts)],
                   check=True, timeout=120)
    raw = np.fromfile(out, dtype=np.uint8)
    f = raw[: 9 * 4 * n].view(np.float32).reshape(9, n)
    ids = raw[9 * 4 * n:].view(np.int64)
    x0, y0, z0, x, y, z, vx, vy, vz = f
    tp = O.TreeParams(np.zeros(3, np.float32), np.full(3, box, np.float32),
                      np.full(3, 1.0, np.float32), np.full(3, box - 1.0, np.float32),
                      fsrrmax=float(np.float32(3.116326355)), rsm=0.1, oa=0.1, nd=nd,
                      tdpts=tdpts)
    tree = O.build_tree(tp, x0, y0, z0)
    assert np.array_equal(ids, tree.id)
    assert np.array_equal(x.view(np.uint32), tree.x.view(np.uint32))
    lists = O.build_lists(tp, tree)
    rx, ry, rz = O.forces_step16(tp, tree, lists)
    V = np.stack([vx, vy, vz], 1)
    Rv = fcoeff * np.stack([rx, ry, rz], 1)
    err = np.sqrt(np.sum((V - Rv) ** 2) / np.sum(Rv ** 2))
    assert err <= 1e-5, err
