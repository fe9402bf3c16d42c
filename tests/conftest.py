import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: longer CPU-side oracle runs")


@pytest.fixture(scope="session")
def ctx():
    from ssf_amd import gpu
    c = gpu.Context(0)
    yield c
    # matrices hold a reference to the context; it is destroyed after the last one


@pytest.fixture(autouse=True)
def _name_the_case(request):
    # print the case before it runs so a failing GPU step is identifiable
    print(f"\n[case] {request.node.nodeid}", flush=True)
    yield
