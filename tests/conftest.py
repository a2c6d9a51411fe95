import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def _gpu_ok():
    try:
        import fastpls_amd
        return fastpls_amd.available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def ctx():
    # torch ships its own HIP runtime, which only initialises before the
    # library's (tests hand torch device tensors to the library)
    import torch
    if torch.cuda.is_available():
        torch.cuda.init()
    import fastpls_amd
    if not _gpu_ok():
        pytest.fail("GPU test requested but no gfx950 device / library is available")
    return fastpls_amd.context(0)
