import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# pytest-xdist workers share the host's cores: give each worker's OpenMP
# (the CPU oracle) its share instead of one thread per core in every worker
if os.environ.get("PYTEST_XDIST_WORKER") and "OMP_NUM_THREADS" not in os.environ:
    n = int(os.environ.get("PYTEST_XDIST_WORKER_COUNT", "1"))
    os.environ["OMP_NUM_THREADS"] = str(max(1, (os.cpu_count() or 1) // max(1, n)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(autouse=True)
def _release_device_memory(request):
    """After each GPU test, drop unreferenced objects and torch's cached
    blocks, so that one test's propagators (tens of GB of checkpoints on the
    full-size workloads) never starve the next."""
    yield
    if request.node.get_closest_marker("gpu") is None:
        return
    import gc
    gc.collect()
    torch = sys.modules.get("torch")
    if torch is not None and torch.cuda.is_available():
        torch.cuda.empty_cache()
