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


def pytest_terminal_summary(terminalreporter):
    """The tie log: every label the parity tests excused as a near-tie."""
    try:
        from tests.helpers import TIE_LOG
    except Exception:
        return
    if not TIE_LOG:
        return
    terminalreporter.write_sep("-", f"tie log: {len(TIE_LOG)} excused near-tie rows")
    for e in TIE_LOG:
        terminalreporter.write_line(
            f"{e['test']} row {e['row']}: oracle margin {e['oracle_margin']:.3e}, "
            f"device margin {e['gpu_margin'] if e['gpu_margin'] is None else format(e['gpu_margin'], '.3e')}, "
            f"bound {e['bound']:.3e}")
    path = os.environ.get("FPLS_TIE_LOG")
    if path:
        import json
        with open(path, "a") as f:
            for e in TIE_LOG:
                f.write(json.dumps(e) + "\n")
