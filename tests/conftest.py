import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(REPO, "gdistger-re_amd"), os.path.join(REPO, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


@pytest.fixture(scope="session")
def c1_edges():
    """Chung-Lu, 10,000 vertices, ~50k edges, gamma 2.5, min degree 2, seed 0."""
    import distger
    return distger.gen_chung_lu(10000, 50000, 2.5, 2, 0)


@pytest.fixture(scope="session")
def ctx():
    import distger
    c = distger.Context(0)
    yield c
    c.close()
