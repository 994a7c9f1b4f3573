import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(Path(__file__).resolve().parent))


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (ROCm) GPU and the built HIP library')
    config.addinivalue_line('markers', 'slow: long CPU test')


@pytest.fixture(scope='session')
def built_lib():
    """Path of libsawfem_hip.so, building it in-tree if it is missing."""
    from sawsim_amd import _lib
    if not _lib.LIB_PATH.exists():
        from sawsim_amd.build import build
        build()
    return _lib.LIB_PATH
