"""Shared test configuration.

Tests marked ``gpu`` run the HIP engine and need an MI355X; the others check
the oracle (the CPU restatement of the reference under ``oracle/``) against
the reference's own golden values, and the host logic, on the CPU.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (runs the HIP engine)")
    config.addinivalue_line("markers", "slow: long-running")


def _gpu_available():
    try:
        from akitsushima_amd import _lib
        return _lib.device_count() > 0
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if _gpu_available():
        return
    skip = pytest.mark.skip(reason="no GPU visible")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)
