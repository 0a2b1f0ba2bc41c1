import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libnslx.so on cuda:0)")


def _ensure_lib():
    lib = os.path.join(ROOT, "nsl_amd", "libnslx.so")
    if not os.path.exists(lib):
        from nsl_amd import build
        build.build()
    return lib


@pytest.fixture(scope="session")
def libpath():
    return _ensure_lib()


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    _ensure_lib()
    import nsl_amd
    return nsl_amd
