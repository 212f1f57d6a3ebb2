import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (REPO, os.path.join(REPO, "olf-gpu_amd"), os.path.join(REPO, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI library)")


_NETS = {}


@pytest.fixture(scope="session")
def grid():
    """grid(name) -> cached gridgen.Network of a named workload."""
    from olfgpu import gridgen

    def get(name, load_scale=None):
        key = (name, load_scale)
        if key not in _NETS:
            _NETS[key] = gridgen.workload(name, load_scale)
        return _NETS[key]

    return get
