import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    config.addinivalue_line("markers", "slow: long-running CPU test")

# import torch once at collection time: the first import on a fresh host can
# take minutes and must not count against a test's time limit
try:
    import torch  # noqa: F401
except Exception:  # pragma: no cover
    pass
