import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def H():
    import hvvr_amd

    hvvr_amd.load_host()
    return hvvr_amd


@pytest.fixture(scope="session")
def O():
    import pyoracle

    pyoracle.lib()
    return pyoracle


@pytest.fixture(scope="session")
def scene1(H):
    return H.HostScene(H.SCENE_HEIGHTFIELD_BOXES, 0)


@pytest.fixture(scope="session")
def view1(H):
    return H.HostView(H.VIEW_PRESET_HEIGHTFIELD)


@pytest.fixture(scope="session")
def view1_small(H):
    return H.HostView(H.VIEW_PRESET_HEIGHTFIELD, 0, 160, 120)


@pytest.fixture(scope="session")
def scene2(H):
    return H.HostScene(H.SCENE_ROOM, 0, 1024)


@pytest.fixture(scope="session")
def view2_small(H):
    return H.HostView(H.VIEW_PRESET_ROOM_EYE, 0, 270, 300)


@pytest.fixture(scope="session")
def ctx(H):
    c = H.GpuContext(0)
    yield c
    c.close()


def rgba_delta(a, b):
    return np.abs(np.asarray(a, np.uint32).view(np.uint8).astype(np.int32) -
                  np.asarray(b, np.uint32).view(np.uint8).astype(np.int32))
