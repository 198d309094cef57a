import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU and the built HIP library")
    config.addinivalue_line("markers", "default_exact_order: run with the library's default "
                            "exact-plane order (the Wiener planes by member pairs) and "
                            "the Wiener member transforms on the matrix cores")


@pytest.fixture(autouse=True)
def _serial_exact_order(request):
    """GPU tests compare bit for bit with the oracle, which sums the exact
    planes and the member transforms in the reference's serial order: they
    run both stages in that order unless marked default_exact_order (those
    hold the pair order and the matrix-core transforms to their own
    tolerance)."""
    if request.node.get_closest_marker("gpu") is None or \
            request.node.get_closest_marker("default_exact_order") is not None:
        yield
        return
    from bmnd_amd import engine
    prev = engine.set_exact_order(ht="serial", wiener="serial", transforms="serial")
    try:
        yield
    finally:
        engine.set_exact_order(**prev)


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test scheduled on a machine without a GPU")
    return torch.device("cuda", 0)
