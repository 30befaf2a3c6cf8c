import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: large workload")


@pytest.fixture(scope="session")
def renderer_factory():
    """Creates product renderers (HIP) and closes them at the end of the session."""
    from pathfinder_amd import DestFramebuffer, Renderer, RendererOptions
    made = []

    def make(width, height, background, **kw):
        r = Renderer(RendererOptions(DestFramebuffer.full_window(width, height), background),
                     **kw)
        made.append(r)
        return r

    yield make
    for r in made:
        r.close()
