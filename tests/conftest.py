import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long-running CPU test")


# HD_EMU=1 runs the GPU tests against the CPU-emulated build of the library
# (tests/emu): every kernel on the host, device buffers against guard pages.
if os.environ.get("HD_EMU") == "1":
    import emu_lib

    emu_lib.use_emulator()
