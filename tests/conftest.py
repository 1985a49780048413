import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: larger meshes")


@pytest.fixture(scope="session")
def c0():
    from fesom2_amd import meshgen as mg
    mesh, state, steps = mg.make_config("c0")
    return mesh, state, steps


@pytest.fixture(scope="session")
def c0_oracle(c0):
    from oracle.oracle import Oracle
    return Oracle(c0[0])
