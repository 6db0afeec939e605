import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI of libquantit_amd.so)")


@pytest.fixture(scope="session")
def ctx():
    import quantit_amd as qa
    c = qa.Context(device=0, arena_bytes=8 << 30)
    yield c
    c.close()
