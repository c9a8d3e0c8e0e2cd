import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def golden_sfmt():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "sfmt_vectors.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def gpu():
    """Fails loudly when a GPU test runs where the HIP library sees no device."""
    from cerebellum_simulation_amd import hip_lib
    n = hip_lib().cbm_device_count()
    assert n > 0, "no HIP device visible to libcbm_hip.so"
    return n
