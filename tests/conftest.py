import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (runs the HIP library)')


def pytest_collection_modifyitems(config, items):
    # print the test id before each GPU test so a failing run names its case
    pass


@pytest.fixture(autouse=True)
def _announce(request):
    if request.node.get_closest_marker('gpu'):
        print(f'\n[gpu-test] {request.node.nodeid}', flush=True)
    yield
