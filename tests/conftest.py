import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: test needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: longer CPU test")
    # the oracle's OpenMP loops over cells: up to 16 host threads (the GPU box's share)
    import oracle as O
    O.set_threads(min(16, os.cpu_count() or 1))
