import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); runs through the C-ABI library")
    config.addinivalue_line("markers", "slow: longer CPU runs")


def pytest_collection_modifyitems(config, items):
    """A gpu test on a machine without a HIP device is skipped, as the reference's conftest does for CuPy."""
    if os.environ.get("WADI2D_FORCE_GPU_TESTS"):
        return
    try:
        from wadi2d_amd.solver.hip_backend import device_count

        have = device_count() > 0
    except Exception as error:
        if os.path.exists("/dev/kfd"):  # a GPU machine whose library will not load: say so, never skip
            raise pytest.UsageError(f"the HIP library does not load on a GPU machine: {error}") from error
        have = False
    if have:
        return
    skip = pytest.mark.skip(reason="no HIP device")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)
