"""Test setup: the `gpu` marker, import paths for the package (its directory name has a hyphen, so it is
imported with importlib) and the oracle."""

from __future__ import annotations

import importlib
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.dirname(os.path.abspath(__file__))):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on the device)")


@pytest.fixture(scope="session")
def lw():
    return importlib.import_module("livox-warp_amd")


@pytest.fixture(scope="session")
def orc():
    import oracle

    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def gpu_ok():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU test collected on a machine without a visible GPU")
    return True
