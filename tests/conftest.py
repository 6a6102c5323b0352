import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libketsv on the device)")


@pytest.fixture(autouse=True)
def _default_options():
    import ket_amd
    old = ket_amd.configure()
    yield
    ket_amd.configure(**old)
