import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


@pytest.fixture(scope="session")
def small_case():
    """Diamond, 1st-shell clusters, 4x4x4 mesh: the oracle finishes in seconds."""
    from fourphonon_amd import crystal
    return crystal.config_case(1, ngrid=(4, 4, 4))


@pytest.fixture(scope="session")
def small_oracle(small_case):
    from oracle.oracle import Oracle
    return Oracle(small_case)
