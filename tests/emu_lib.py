"""Points hydist_c_amd.engine at the CPU-emulated build of the library
(tests/_build/emu/libhydist_emu.so) for bounds-checked kernel runs on the
host.  Test infrastructure only."""
import os

from hydist_c_amd import engine

EMU_LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_build", "emu", "libhydist_emu.so")


def use_emulator():
    """HD_EMU_LIB names another emulated build (a sanitized one) to load
    instead of tests/_build/emu's."""
    other = os.environ.get("HD_EMU_LIB")
    if other:
        engine.LIB_PATH = other
        engine._lib = None
        return engine.lib()
    if not os.path.exists(EMU_LIB):
        import sys
        sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "emu"))
        import build_emu
        build_emu.build()
    engine.LIB_PATH = EMU_LIB
    engine._lib = None
    return engine.lib()
