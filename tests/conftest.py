import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def ctx():
    from femocs_amd import fgpu
    return fgpu.Context(0)


@pytest.fixture(scope="session")
def small_mesh():
    from femocs_amd import mesh as M
    return M.make_tip_mesh(M.small_geometry())
