import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and libeuvl.so")
    config.addinivalue_line("markers", "slow: long CPU run (minutes); skipped unless EUVL_SLOW=1")


def pytest_collection_modifyitems(config, items):
    if os.environ.get("EUVL_SLOW") == "1":
        return
    skip = pytest.mark.skip(reason="slow oracle run; set EUVL_SLOW=1")
    for it in items:
        if "slow" in it.keywords:
            it.add_marker(skip)


@pytest.fixture(scope="session")
def golden():
    return GOLDEN
