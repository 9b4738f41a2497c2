import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
for p in (ROOT / "luce-render_amd", ROOT):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


@pytest.fixture(scope="session")
def device():
    """The GPU for -m gpu tests; the library must be built (no fallback)."""
    from luce_render import render
    dev = render.Device.open()
    yield dev
    dev.destroy()


@pytest.fixture(scope="session")
def cornell_probe():
    """The probe workload: Cornell box + 10,000 random triangles, compiled."""
    from luce_render import render, scenes
    return render.compile_scene(scenes.cornell_probe_set())
