import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "osprey-dpd_amd"), os.path.join(ROOT, "oracle"), os.path.dirname(__file__)):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


def pytest_collection_modifyitems(config, items):
    # GPU tests run in one process on the GPU machine; elsewhere they are skipped
    # only when deselected with -m "not gpu".
    pass
