import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def _gpu_visible():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    # GPU tests are skipped only when no device is visible at all; on a GPU
    # box a missing libnpbx.so makes them fail, never skip.
    if _gpu_visible():
        return
    skip = pytest.mark.skip(reason="no HIP device visible")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


@pytest.fixture(scope="session", autouse=True)
def _native_library_built():
    """Builds libnpbx.so once if it is absent (hipcc cross-compiles on CPU)."""
    from npbench_amd import _native
    if not _native.LIB_PATH.exists():
        from npbench_amd import build
        build.build(verbose=False)
