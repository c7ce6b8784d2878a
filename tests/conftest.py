import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and libfvm_gpu.so")


def _stamp(msg):
    sys.stdout.write(msg + "\n")
    sys.stdout.flush()


@pytest.fixture(autouse=True)
def _name_the_case(request):
    # print the case before it runs so a GPU failure can be attributed
    if request.node.get_closest_marker("gpu") is not None:
        _stamp(f"[gpu case] {request.node.nodeid}")
    yield
