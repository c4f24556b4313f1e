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
