import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libkmst on cuda:0)")


@pytest.fixture(scope="session")
def gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import kamsta_amd
    kamsta_amd.load()
    return torch


@pytest.fixture(scope="session")
def ctx(gpu):
    import kamsta_amd
    c = kamsta_amd.Context(0)
    yield c
    c.close()
