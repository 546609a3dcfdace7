"""Shared pytest setup: the `gpu` marker, import paths and built libraries.

Tests marked `gpu` need an MI355X and the built HIP library; everything else
runs on the CPU.  The CPU oracle (oracle/) is test infrastructure: tests load
it as the checker, never as the thing under test.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an AMD Instinct GPU (gfx950) and the built HIP library")


def _ensure_oracle():
    from oracle import pyoracle
    if not os.path.exists(pyoracle.LIB_PATH):
        subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle")])


def _ensure_product():
    import superconga_amd
    if not os.path.exists(superconga_amd.LIB_PATH):
        from superconga_amd import build
        build.build()


@pytest.fixture(scope="session")
def oracle():
    """The CPU oracle module (builds oracle/_build if it is missing)."""
    _ensure_oracle()
    from oracle import pyoracle
    return pyoracle


@pytest.fixture(scope="session")
def product():
    """The superconga_amd package with its library loaded."""
    _ensure_product()
    import superconga_amd
    superconga_amd.lib()
    return superconga_amd
