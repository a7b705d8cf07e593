import os
import sys

import pytest

try:
    # torch's bundled HIP runtime must be loaded before libsky_hip.so pulls in
    # /opt/rocm's, or torch later sees no device.
    import torch  # noqa: F401
except ImportError:
    pass

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "sky-tracer_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP integrator)")
    config.addinivalue_line("markers", "slow: longer CPU statistical test")


@pytest.fixture(scope="session")
def gpu_context():
    import pysky

    print(f"[gpu] creating context on device 0 ({pysky.lib().sky_version().decode()})", flush=True)
    return pysky.Context(0)
