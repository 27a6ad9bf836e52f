import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and libtcw_hip")


@pytest.fixture(scope="session")
def hip_ready():
    """Initialise the "hip" map version once; fail loudly if it cannot run."""
    from pyfstat_amd import tcw_fstat_map_funcs as tcw

    features, ctx = tcw.init_transient_fstat_map_features("hip")
    assert features["hip"] and ctx is None
    return features
