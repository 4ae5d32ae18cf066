"""CPU checks of the C ABI: the library loads and exports every entry point
declared in include/stda_hip.h (no compute without a GPU)."""
import ctypes
import os
import re

from conftest import ROOT


def declared_functions():
    src = open(os.path.join(ROOT, "include", "stda_hip.h")).read()
    return sorted(set(re.findall(r"\b(stda_hip_[a-z_]+)\s*\(", src)))


def test_header_matches_python_table():
    from stda_amd.core import C_API
    assert declared_functions() == sorted(C_API)


def test_library_exports_every_declared_symbol():
    from stda_amd.core import LIB_PATH, load_library
    assert os.path.exists(LIB_PATH), "build libstda_hip.so first (python -m stda_amd.build)"
    L = load_library()
    for name in declared_functions():
        assert hasattr(L, name), name


def test_error_strings_without_device():
    from stda_amd.core import load_library
    L = load_library()
    assert L.stda_hip_error_string(2).decode().startswith("no CSF")
    assert L.stda_hip_error_string(7).decode() == "matrix power impossible"


def test_create_rejects_null():
    from stda_amd.core import load_library
    L = load_library()
    assert L.stda_hip_create(None, 0) == 1
