import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(autouse=True)
def _fresh_bc_registry():
    """Boundary ids are process-global and uint8-limited; restart them for every test."""
    from xlb_amd.operator.boundary_condition import boundary_condition_registry

    boundary_condition_registry.reset()
    yield
