import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built libhgx.so")


@pytest.fixture(scope="session")
def ctx():
    import hypergraphrewritingengine_amd as hgx
    c = hgx.Context(0)
    print("\n[device]", c.info, flush=True)
    yield c
