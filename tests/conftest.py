import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


# BB_EMU=1 points the loader at the CPU emulation build of the HIP library
# (tests/emu, AddressSanitizer), so the GPU tests can run on the host to
# check kernel indexing before anything runs on a GPU.
EMU = os.environ.get("BB_EMU") == "1"
if EMU:
    from bluebottle_amd import runtime as _rt
    _emu_lib = os.path.join(ROOT, "tests", "emu", "_build", "lib")
    _rt.HIP_LIB = os.path.join(_emu_lib, "libbluebottle_hip.so")
    _rt.HOST_LIB = os.path.join(_emu_lib, "libbluebottle_host.so")

def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import pyoracle
    pyoracle.build()
    return pyoracle.load()
