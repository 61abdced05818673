import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); runs through the C ABI")


def pytest_collection_modifyitems(config, items):
    # run GPU tests only where a HIP device exists; elsewhere mark them skipped
    if os.environ.get("RAV2D_FORCE_GPU_TESTS"):
        return
    try:
        import rav2d_amd
        have = rav2d_amd.device_count() > 0
    except Exception:
        have = False
    if have:
        return
    skip = pytest.mark.skip(reason="no HIP device in this environment")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)
