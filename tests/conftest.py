import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "webgpu-ocean_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def _say(msg):
    sys.stdout.write(msg + "\n")
    sys.stdout.flush()


@pytest.fixture(autouse=True)
def _announce(request):
    # Names the running case before any GPU work so a crash log shows it.
    if request.node.get_closest_marker("gpu"):
        _say(f"\n[gpu-case] {request.node.nodeid}")
    yield
