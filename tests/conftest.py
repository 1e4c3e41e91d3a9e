import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def _ensure_built():
    """Build the oracle and the HIP library in-tree if they are missing."""
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle_mgpcg.so")):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True)
    if not os.path.exists(os.path.join(ROOT, "cirrus_amd", "libcirrus_mgpcg.so")):
        subprocess.run([sys.executable, "-m", "cirrus_amd.build", "-q"], check=True, cwd=ROOT)


_ensure_built()


def gpu_available():
    if not os.path.exists("/dev/kfd"):
        return False
    import cirrus_amd
    return cirrus_amd.device_count() > 0


def pytest_collection_modifyitems(config, items):
    if gpu_available():
        return
    skip = pytest.mark.skip(reason="no GPU in this environment")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)
