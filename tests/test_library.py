"""The C-ABI library loads and exports every symbol include/hydist.h
declares (no GPU calls)."""
import ctypes as C
import os
import re

import pytest

from hydist_c_amd import abi, engine

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hd_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = engine.lib()
    names = declared("hydist.h")
    assert len(names) > 15
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_library_exports_the_reference_host_surface():
    # include/hydist_compat.h: the C++ functions of engine.h / loader.h
    import subprocess
    out = subprocess.run(["nm", "-DC", "--defined-only", engine.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    text = open(os.path.join(ROOT, "include", "hydist_compat.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"^\s*(?:void|int|Argument_Pointers|Array_Pointers|hd_ctx \*|"
                           r"std::pair<int, int>)\s+\*?([A-Za-z_][A-Za-z0-9_]*)\s*\(", text, flags=re.M))
    assert {"Hydraulic_Calculation", "load_inputs", "Load_coeffs", "save_result"} <= names
    missing = [n for n in names if f" {n}(" not in out and f" {n}<" not in out]
    assert not missing, missing


def test_struct_sizes_match_reference_layout():
    # Constant_Coeffs: 49 doubles; Options: 4 doubles, int, 2 bools, 2 ints, 4 bools, int
    assert C.sizeof(abi.Coeffs) == 49 * 8
    assert C.sizeof(abi.Options) == 56


def test_load_coeffs_matches_oracle():
    import oracle_lib as ol
    a = engine.load_coeffs()
    b = ol.load_coeffs()
    assert bytes(a) == bytes(b)
    a = engine.load_coeffs({"dm": 2e-4, "Toe": 0.1})
    b = ol.load_coeffs({"dm": 2e-4, "Toe": 0.1})
    assert bytes(a) == bytes(b)


def test_create_without_device_fails_cleanly():
    rc, info = engine.device_info()
    if rc == 0:
        pytest.skip("a device is present")
    from hydist_c_amd import scenarios as S
    with pytest.raises(engine.HydistError):
        engine.Simulation(S.open_box(20, 16))
