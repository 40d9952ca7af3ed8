"""Shared fixtures.  Tests marked ``gpu`` need an MI355X (run with ``-m gpu``); the rest
run on the CPU only: the oracle against the golden vectors, host logic, and that the
C-ABI library loads and exports what include/qsvh.h declares."""
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden", "golden.npz")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def golden():
    z = np.load(GOLDEN, allow_pickle=False)
    meta = json.loads(bytes(z["meta"]).decode())
    return z, meta


@pytest.fixture(scope="session")
def backend():
    import rustatevec_amd as rs
    b = rs.HipBackend(0)
    yield b
    b.close()
