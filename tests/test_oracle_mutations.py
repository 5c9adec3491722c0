"""The pins of the oracle's walk catch plausible mistakes.

Each case plants one mistake in a copy of oracle/hsr_oracle.c (a text
substitution that must match exactly once), compiles the copy into a
temporary library, points the oracle's wrappers at it and runs the
independent checks of tests/oracle_pins.py: at least one must fail.  The
real oracle passes them all (tests/test_oracle_lists.py).  CPU only.
"""
import os
import shutil
import subprocess

import pytest

import oracle_pins as P
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "hsr_oracle.c")

# (name, text in hsr_oracle.c, replacement)
MUTATIONS = [
    ("children pushed right then left (left subtree walked first)",
     "int64_t ch[2] = {t[tln].cl, t[tln].cr};",
     "int64_t ch[2] = {t[tln].cr, t[tln].cl};"),
    ("ancestor's children pushed right then left",
     "        if (tlncl != tl && tlncl > 0 && t[tlncl].count > 0) q[nq++] = tlncl;\n"
     "        if (tlncr != tl && tlncr > 0 && t[tlncr].count > 0) q[nq++] = tlncr;",
     "        if (tlncr != tl && tlncr > 0 && t[tlncr].count > 0) q[nq++] = tlncr;\n"
     "        if (tlncl != tl && tlncl > 0 && t[tlncl].count > 0) q[nq++] = tlncl;"),
    ("opening test on the diagonals uses >=",
     "if (dh2 > dtt2) {", "if (dh2 >= dtt2) {"),
    ("opening test on the shortest edge uses >=",
     "if (l2 > dtt2) {", "if (l2 >= dtt2) {"),
    ("far cut uses >=",
     "if (dist2 > fsrrmax2) continue;", "if (dist2 >= fsrrmax2) continue;"),
    ("four corner diagonals dropped",
     "for (int i = 0; i < 2 && !looksBig; ++i)", "for (int i = 0; i < 0 && !looksBig; ++i)"),
    ("diagonal with z taken max -> max",
     "float z1 = t[tln].xmin[2] - t[tl].xc[2];", "float z1 = t[tln].xmax[2] - t[tl].xc[2];"),
    ("L-infinity gap test uses >=",
     "if (dist > fsrrmax) {", "if (dist >= fsrrmax) {"),
    ("force-box inside test on closed intervals",
     "inside &= (t[tl].xmax[i] < p->max_force[i] &&\n"
     "                   t[tl].xmax[i] > p->min_force[i]) ||\n"
     "                  (t[tl].xmin[i] < p->max_force[i] &&\n"
     "                   t[tl].xmin[i] > p->min_force[i]);",
     "inside &= (t[tl].xmax[i] <= p->max_force[i] &&\n"
     "                   t[tl].xmax[i] >= p->min_force[i]) ||\n"
     "                  (t[tl].xmin[i] <= p->max_force[i] &&\n"
     "                   t[tl].xmin[i] >= p->min_force[i]);"),
    ("leaf order left-first",
     "      if (t[tl].cl > 0) q[nq++] = t[tl].cl;\n      if (t[tl].cr > 0) q[nq++] = t[tl].cr;",
     "      if (t[tl].cr > 0) q[nq++] = t[tl].cr;\n      if (t[tl].cl > 0) q[nq++] = t[tl].cl;"),
    ("pptdr takes the largest face distance",
     "  float r = (f < e) ? f : e;\n  r = (r < d) ? r : d;\n  r = (r < c) ? r : c;\n"
     "  r = (r < b) ? r : b;\n  r = (r < a) ? r : a;",
     "  float r = (f > e) ? f : e;\n  r = (r > d) ? r : d;\n  r = (r > c) ? r : c;\n"
     "  r = (r > b) ? r : b;\n  r = (r > a) ? r : a;"),
    ("pptdr drops one face",
     "  r = (r < a) ? r : a;\n  return r;", "  return r;"),
    ("P5 coefficient a4 one digit off",
     "0.0000605491f", "0.0000605492f"),
    ("P6 coefficient a2 one digit off",
     "0.0133122f", "0.0133123f"),
]


@pytest.fixture(scope="module")
def workdir(tmp_path_factory):
    if shutil.which("gcc") is None:
        pytest.skip("gcc is not available")
    return tmp_path_factory.mktemp("mutants")


@pytest.mark.parametrize("name,old,new", MUTATIONS, ids=[m[0] for m in MUTATIONS])
def test_planted_mistake_fails_a_pin(workdir, name, old, new):
    src = open(SRC).read()
    assert src.count(old) == 1, f"mutation site not unique or missing: {name}"
    k = MUTATIONS.index((name, old, new))
    path_c = os.path.join(workdir, f"mut{k}.c")
    path_so = os.path.join(workdir, f"libmut{k}.so")
    with open(path_c, "w") as f:
        f.write(src.replace(old, new))
    subprocess.check_call(["gcc", "-O2", "-fPIC", "-ffp-contract=off", "-fopenmp", "-std=c11",
                           "-w", "-I" + os.path.join(ROOT, "oracle"), "-shared", "-o", path_so,
                           path_c, "-lm"])
    prev = O.use(O.load(path_so))
    try:
        failed = P.run_all(P.tie_cases() + P.generic_cases())
    finally:
        O.use(prev)
    assert failed, f"no pin fails when the oracle has this mistake: {name}"
