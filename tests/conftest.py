import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "cellbasedmodels.jl_amd"), os.path.join(ROOT, "oracle"), os.path.dirname(__file__)):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libcbm_hip.so)")


@pytest.fixture(scope="session")
def gpu_available():
    import torch  # device discovery only
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return True
