import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "stochastic-density-functional-theory-piny_amd")
for p in (REPO, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU and the built HIP library")


@pytest.fixture(scope="session")
def tiny_system():
    from sdft_amd import systems
    # Si8 cell, 3 bohr^-2 sphere on a 24^3 grid (ncoef ~ 50): small enough for
    # dense-H references
    return systems.build_system("si8-tiny", 10.26, 1, 24, g2max=3.0)


@pytest.fixture(scope="session")
def si8_system():
    from sdft_amd import systems
    return systems.build_system("si8", 10.26, 1, 32)
