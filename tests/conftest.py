from __future__ import annotations

import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config) -> None:
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


def _device_count() -> int:
    try:
        from mc_dagprop_amd import capi

        return capi.device_count()
    except Exception:
        return 0


@pytest.fixture(scope="session")
def gpu() -> None:
    """Fails (not skips) a gpu-marked test when no HIP device is visible: a GPU test must run on the GPU."""
    if _device_count() < 1:
        pytest.fail("no HIP device visible to libmcdp")


@pytest.fixture(scope="session")
def has_gpu() -> bool:
    return _device_count() > 0
