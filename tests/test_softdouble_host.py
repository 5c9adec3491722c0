"""The exact double arithmetic the tree build uses on the device
(hacc-coral_amd/csrc/softdouble.h: serial additions, the exact integer-sum
path, fixed-grid maps and the double-to-float quotient of BGQCM.c:208-210),
compiled for the host and checked against native IEEE double arithmetic by
tests/cpp/test_softdouble.cpp."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_softdouble_matches_native_double(tmp_path):
    exe = tmp_path / "test_softdouble"
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-o", str(exe),
                           os.path.join(ROOT, "tests", "cpp", "test_softdouble.cpp")])
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    fields = out.stdout.split()
    # the exact path must take almost every trial (unit masses and ties included)
    assert int(fields[fields.index("exact") + 1]) > 19000, out.stdout
