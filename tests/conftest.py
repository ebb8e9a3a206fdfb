import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")


def _have_gpu() -> bool:
    # the AMD kernel driver node; present on a GPU box, absent on CPU-only hosts
    return os.path.exists("/dev/kfd")


def pytest_collection_modifyitems(config, items):
    if _have_gpu():
        return
    skip = pytest.mark.skip(reason="no GPU in this environment")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)


@pytest.fixture(autouse=True)
def _fresh_tuning_table(request):
    """Every GPU test starts from an empty launch-shape table (the fixed
    rule), whatever earlier tests loaded; tests that want the committed
    table load it themselves."""
    if "gpu" in request.keywords and _have_gpu():
        from ples_amd import mg

        mg.tuning_clear()
    yield
