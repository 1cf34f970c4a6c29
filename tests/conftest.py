import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

SCENES = sorted((ROOT / "tests" / "golden" / "scenes").glob("scene*.json"), key=lambda p: int(p.stem[5:]))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


def _have_gpu() -> bool:
    try:
        import pathtracer_amd as pt

        ctx = pt.Context(0)
        ctx.close()
        return True
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu_ctx():
    import pathtracer_amd as pt

    ctx = pt.Context(0)
    yield ctx
    ctx.close()


@pytest.fixture(scope="session")
def scene_texts():
    return {p.stem: p.read_text() for p in SCENES}
