import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs a ROCm GPU (MI355X)')
    config.addinivalue_line('markers', 'slow: long-running')


def _gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope='session')
def gpu():
    if not _gpu_available():
        pytest.skip('no GPU')
    import torch
    torch.cuda.init()
    from spinterps_amd import _lib
    _lib.lib()  # fail loudly when the library is missing
    return torch.device('cuda', 0)


def pytest_collection_modifyitems(config, items):
    # Stable order: CPU tests first, GPU tests after.
    items.sort(key=lambda it: 1 if it.get_closest_marker('gpu') else 0)
