import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: runs the HIP path on an MI355X")


@pytest.fixture(scope="session")
def pyoracle():
    import pyoracle as po
    if not os.path.exists(po.LIB_PATH):
        po.build()
    return po


def problem(name):
    from citcom_amd import deck
    return deck.problem_from_deck({k: str(v) for k, v in deck.workload(name).items()})


def initial_temperature(name, p):
    """The starting temperature of a deck: the reference's shipped restart
    field for busse2 (examples/Busse1993/case2-tic.temp.0.0, as its restart
    reader leaves it), else the perturbed conductive profile."""
    from citcom_amd import deck
    if name.split("@")[0] == "busse2":
        return golden("busse2_tic.npz")["T"]
    return deck.perturbed_temperature(p)


def golden(name):
    return np.load(os.path.join(GOLDEN, name))


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
