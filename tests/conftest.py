"""Shared pytest setup: markers, import paths, GPU gating."""
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT / "tesina-rt_amd"))
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built librtx.so")
    config.addinivalue_line("markers", "slow: larger CPU oracle runs")


def gpu_available() -> bool:
    try:
        import torch  # noqa: F401  (device count only; no GPU initialisation)
        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def ctx():
    """One rtx context for the whole GPU session (the library, never a fallback)."""
    import rtx
    c = rtx.Context(0)
    yield c
    c.close()
