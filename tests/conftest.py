import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "dirac-research_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    import __graft_entry__

    __graft_entry__._ensure_built()


@pytest.fixture(scope="session")
def ctx():
    import dirac_me

    c = dirac_me.Context(0)
    yield c
    c.close()
