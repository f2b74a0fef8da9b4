import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "sparse-blas_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libspx.so on the device)")


def pytest_collection_modifyitems(config, items):
    # GPU tests are skipped cleanly where no GPU is visible; -m gpu selects them.
    try:
        import spx
        have_gpu = os.path.exists(spx.LIB_PATH) and spx.lib().spx_device_count() > 0
    except Exception:
        have_gpu = False
    if have_gpu:
        return
    skip = pytest.mark.skip(reason="no GPU visible (libspx.so reports 0 devices)")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)
