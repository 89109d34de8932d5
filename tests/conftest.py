import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import pyoracle
    if not pyoracle.available():
        pytest.skip("oracle/liboracle.so not built (run __graft_entry__.build())")
    return pyoracle


@pytest.fixture(scope="session")
def ref_lib():
    from oracle import refshim
    if not refshim.available():
        pytest.skip("oracle/_ref not built (reference sources absent here)")
    return refshim
