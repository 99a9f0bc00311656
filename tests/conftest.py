"""Test configuration: the ``gpu`` marker and import paths.

Tests marked ``gpu`` call libmivegas.so on a HIP device and are run with
``-m gpu`` on an MI355X; everything else runs on the CPU.
"""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and libmivegas.so")
