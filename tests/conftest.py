import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
for p in (str(ROOT), str(ROOT / "wgpu-cpu_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libwgc on the GPU)")
    # the oracle is plain C; make rebuilds it only when its sources changed
    # (on a machine without a compiler the prebuilt library is used as is)
    lib = ROOT / "oracle" / "liboracle.so"
    try:
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)
    except (OSError, subprocess.CalledProcessError):
        if not lib.exists():
            raise


@pytest.fixture(scope="session")
def gpu_device():
    import wgc
    return wgc.Device(0, wgc.SEMANTICS_REFERENCE)


@pytest.fixture(scope="session")
def gpu_device_ext():
    import wgc
    return wgc.Device(0, wgc.SEMANTICS_EXTENDED)
