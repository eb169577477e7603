"""Shared pytest configuration.

Markers:
  gpu   needs an MI355X (runs the HIP kernels through the C ABI)
  slow  long GPU cases (self-test ranges at p ~ 1e12-2e13)
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (HIP device)")
    config.addinivalue_line("markers", "slow: long-running GPU case")


@pytest.fixture(scope="session")
def golden():
    import json

    def load(name):
        with open(os.path.join(GOLDEN, name)) as f:
            return json.load(f)
    return load


@pytest.fixture(scope="session")
def dev():
    """One GPU context for all GPU tests (fails loudly if the library or the
    GPU is missing: GPU tests never fall back to a CPU path)."""
    import clwilson_amd as cw
    d = cw.Device(0)
    yield d
    d.close()
