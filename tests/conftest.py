import os
import sys

import pytest

# torch (whose ROCm build ships its own HIP runtime) is loaded before the
# library's first call, as a caller that shares streams with torch must do:
# both then use one runtime in the process, and torch stream handles are
# valid for the library.
try:
    import torch  # noqa: F401
except Exception:
    pass

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def oracle():
    import svmtest
    return svmtest.oracle_lib()
