"""Shared pytest configuration.

Tests marked ``gpu`` need an MI355X (gfx950) and the in-tree HIP library;
everything else runs on the CPU: the FP64 oracle against closed forms and
the reference's known-answer vectors, the host-side contract of the C ABI and
the exported symbol set.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an AMD GPU and the built HIP library")
