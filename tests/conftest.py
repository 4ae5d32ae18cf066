import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def water():
    from stda_amd.molden import read_molden
    return read_molden(os.path.join(GOLDEN, "water", "water_sto3g.molden"))


def read_datxy(path):
    rows = []
    with open(path) as fh:
        lines = fh.read().splitlines()
    k = next(i for i, ln in enumerate(lines) if "DATXY" in ln)
    for ln in lines[k + 1:]:
        t = ln.split()
        if t:
            rows.append([float(x) for x in t])
    return rows


def read_table(path):
    with open(path) as fh:
        return [[float(x) for x in ln.split()] for ln in fh if ln.strip()]


def beta_test_omegas():
    """Frequencies of the reference's BETA tests (-resp 2): 0 and the wavelength file's
    1907 and 1064 nm, converted as lresp1 does in single precision (45.56335/lambda,
    linear_response.f:265-268)."""
    import numpy as np
    lam = [float(x) for x in open(os.path.join(GOLDEN, "water", "wavelength")).read().split()]
    return [0.0] + [float(np.float32(np.float32(45.56335) / np.float32(v))) for v in lam]
