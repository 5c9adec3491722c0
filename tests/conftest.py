"""Test configuration: the `gpu` marker and import paths.

CPU tests (`-m "not gpu"`) cover the oracle against the reference's compiled
sources and golden vectors, the host logic, and that libhaccsr.so exports the
C ABI.  GPU tests (`-m gpu`) are the parity tests proper: they call the C ABI
on device 0 and compare with the oracle.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "hacc-coral_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle as O
    if not os.path.exists(O.LIB_PATH):
        O.build()
    return O
