"""Shared test set-up: import paths and the `gpu` marker."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "dynela-v.-4.0_amd")
for p in (PKG, ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libdnlg.so on the GPU)")
    config.addinivalue_line("markers", "slow: CPU test of more than a few seconds")


@pytest.fixture(scope="session")
def ctx():
    import dnlPython as dnl
    c = dnl.Context.default()
    yield c
