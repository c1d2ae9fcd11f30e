import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
ORACLE = os.path.join(ROOT, "oracle")
if ORACLE not in sys.path:
    sys.path.insert(0, ORACLE)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(autouse=True)
def _fresh_tuning_cache():
    """Every test starts with an empty shape-class cache of tuned launch shapes,
    so that a tuning made in one test does not reach the handles of another."""
    try:
        import timstochrigidbodiessim_amd as fa
        fa.clear_tuning_cache()
    except Exception:  # the library is not built: the test reports that itself
        pass
    yield
