import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (ROOT, HERE):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built libspmd.so")
    config.addinivalue_line("markers", "slow: long statistical runs")


def pytest_collection_modifyitems(config, items):
    # print the test name before each GPU test so a crash names its case
    for it in items:
        if it.get_closest_marker("gpu"):
            it.add_marker(pytest.mark.usefixtures("_announce"))


@pytest.fixture
def _announce(request):
    print(f"\n[gpu-test] {request.node.nodeid}", flush=True)
    yield
