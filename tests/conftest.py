"""Shared test setup: the gpu marker, paths, and handles to libcac / the oracle."""
import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "comm-avoiding-cuts-cc_amd"
for p in (str(PKG), str(ROOT / "oracle"), str(ROOT / "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libcac.so on the device)")


@pytest.fixture(scope="session")
def ctx():
    import cac

    c = cac.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def orc():
    import oracle

    return oracle
