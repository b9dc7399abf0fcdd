import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "gpu-rmq_amd"), os.path.join(ROOT, "tests", "golden")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through librmq.so on cuda:0)")
    config.addinivalue_line("markers", "slow: large CPU work (skipped unless RMQ_SLOW=1)")


def pytest_collection_modifyitems(config, items):
    if os.environ.get("RMQ_SLOW") == "1":
        return
    skip = pytest.mark.skip(reason="slow: set RMQ_SLOW=1")
    for item in items:
        if "slow" in item.keywords:
            item.add_marker(skip)


@pytest.fixture(scope="session")
def device():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU test collected but no HIP device is visible")
    return torch.device("cuda:0")
