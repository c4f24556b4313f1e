import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "oxideav-render-vulkan_amd"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP (gfx950) device")


@pytest.fixture(scope="session")
def gpu():
    """A tracer on device 0; GPU tests fail (not skip) without a device."""
    import oxpt

    return oxpt.GpuPathTracer(0)


@pytest.fixture(autouse=True)
def _restore_gpu_state(request):
    """The session tracer is shared: whatever a test changed (kernel form,
    stack split, tail, path cap, environment) is put back after it, pass or
    fail, so one failure cannot change what later tests run."""
    yield
    if "gpu" not in request.fixturenames:
        return
    g = request.getfixturevalue("gpu")
    g.set_tuning(block=256, waves_per_cu=32, max_paths=0xFFFFFFFF)
    g.set_environment(None)
    g.invalidate_scene()
