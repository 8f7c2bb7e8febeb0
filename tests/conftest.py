import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU; runs through the C-ABI")


@pytest.fixture(scope="session")
def ctx():
    """One library context for all GPU tests of the session (fails loudly if
    libstater.so or the GPU is missing: there is no CPU fallback)."""
    import stater_amd

    c = stater_amd.Context(0, timing=False)
    yield c
    c.close()

