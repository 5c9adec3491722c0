"""The oracle's leaf selection, list walk, pseudo-particle radius and law
constants against independent restatements (tests/oracle_pins.py).

The cases include lattices where the walk's float comparisons meet exact
ties (tan_oa 0.5, fsrrmax 2: opening test, far cut and L-infinity gap), leaf
faces lying exactly on the force box's planes, and point-sized degenerate
leaves at tan_oa 0.  tests/test_oracle_mutations.py shows that a mistake
planted in the oracle makes one of these fail.  CPU only.
"""
import numpy as np
import pytest

import oracle_pins as P
from oracle import oracle as O

CASES = P.tie_cases() + P.generic_cases()


@pytest.fixture(scope="module", params=CASES, ids=[c[0] for c in CASES])
def case(request):
    name, tp, x, y, z = request.param
    tree = O.build_tree(tp, x, y, z)
    lists = O.build_lists(tp, tree)
    return tp, tree, lists


def test_leaf_selection_and_order(case):
    tp, tree, _ = case
    P.check_leaf_order(tp, tree)


def test_lists_right_to_left_then_own_leaf(case):
    _, tree, lists = case
    P.check_list_order(tree, lists)


def test_lists_match_recursive_walk(case):
    tp, tree, lists = case
    P.check_list_walk(tp, tree, lists)


def test_pseudo_radius(case):
    tp, tree, _ = case
    P.check_pptdr(tp, tree)


def test_polynomial_constants_are_the_reference_literals():
    P.check_law_constants()


def test_lattice_cases_meet_exact_ties():
    """The tie cases exercise the comparisons at equality: some node's far
    distance equals fsrrmax^2, some diagonal equals dist^2 tan^2 and some
    L-infinity gap equals fsrrmax."""
    hits = {"far": 0, "open": 0, "gap": 0}
    for name, tp, x, y, z in P.tie_cases():
        tree = O.build_tree(tp, x, y, z)
        nodes = tree.nodes
        tan, fsr = np.float32(tp.tan_oa()), np.float32(tp.fsrrmax)
        for tl in O.leaf_order(tp, tree):
            d = nodes["xc"] - nodes["xc"][tl]
            dist2 = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]
            s = nodes["xmax"] - nodes["xmin"]
            diag = s[:, 0] * s[:, 0] + s[:, 1] * s[:, 1] + s[:, 2] * s[:, 2]
            hits["far"] += int(np.sum(dist2 == fsr * fsr))
            hits["open"] += int(np.sum(diag == dist2 * tan * tan))
            lo, hi = nodes["xmin"][tl], nodes["xmax"][tl]
            gap = np.maximum(nodes["xmin"] - hi, lo - nodes["xmax"])
            hits["gap"] += int(np.sum(gap == fsr))
    assert all(v > 0 for v in hits.values()), hits


def test_leaf_faces_on_force_box_planes():
    """In the lattice case the force box's planes hold leaf faces, and the
    leaves whose only faces inside the box lie on its planes are left out."""
    name, tp, x, y, z = P.tie_cases()[0]
    tree = O.build_tree(tp, x, y, z)
    nodes = tree.nodes
    leaf = (nodes["cl"] == 0) & (nodes["cr"] == 0)
    on_plane = np.any((nodes["xmax"] == tp.min_force) | (nodes["xmin"] == tp.max_force), axis=1)
    assert np.any(leaf & on_plane)
    picked = set(O.leaf_order(tp, tree).tolist())
    for k in np.nonzero(leaf & on_plane)[0]:
        inside = np.all(((nodes["xmax"][k] < tp.max_force) & (nodes["xmax"][k] > tp.min_force)) |
                        ((nodes["xmin"][k] < tp.max_force) & (nodes["xmin"][k] > tp.min_force)))
        assert (k in picked) == bool(inside)
