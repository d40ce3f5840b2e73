import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "parallelstencil.jl_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def _has_gpu():
    return os.path.exists("/dev/kfd")


def pytest_collection_modifyitems(config, items):
    if _has_gpu():
        return
    skip = pytest.mark.skip(reason="no GPU in this container")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)


@pytest.fixture(autouse=True)
def _device_errors_stay_with_their_test(request):
    """A GPU test ends with a device synchronisation, so that an error flagged by one of its
    kernels (a suppressed out-of-bounds access) is reported by that test, not by a later one."""
    yield
    if "gpu" in request.keywords and _has_gpu():
        import parallelstencil as ps
        ps.synchronize()
