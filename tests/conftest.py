import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and liblct.so")


# The FDK path's parity files run first and the learned-correction files last,
# so that under -x a failure in the CNN cannot hide a stage or pipeline result.
_FILE_ORDER = (
    "test_gpu_stages.py",
    "test_gpu_pipeline.py",
    "test_gpu_random_geometry.py",
    "test_gpu_large.py",
    "test_gpu_dropin.py",
    "test_gpu_failures.py",
    "test_bench_contract.py",
    "test_gpu_forward.py",
    "test_gpu_cnn.py",
    "test_gpu_learned.py",
)


def _file_rank(item):
    name = os.path.basename(str(item.fspath))
    if name.startswith("test_zz_"):
        return len(_FILE_ORDER) + 1
    try:
        return _FILE_ORDER.index(name)
    except ValueError:
        return len(_FILE_ORDER)


def pytest_collection_modifyitems(session, config, items):
    # stable sort: the order inside a file is kept
    items.sort(key=_file_rank)


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def ctx():
    if not gpu_available():
        pytest.skip("no GPU")
    from listctrecons_amd.recon import Context
    c = Context(0)
    yield c
    c.close()
