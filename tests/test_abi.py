"""CPU checks of the C ABI: the library loads and exports every entry point
declared in include/stda_hip.h (no compute without a GPU)."""
import ctypes
import os
import re

import pytest

from conftest import ROOT


def declared_functions():
    src = open(os.path.join(ROOT, "include", "stda_hip.h")).read()
    return sorted(set(re.findall(r"\b(stda_hip_[a-z_]+)\s*\(", src)))


def test_header_matches_python_table():
    from stda_amd.core import C_API
    assert declared_functions() == sorted(C_API)


@pytest.mark.parametrize("model", ["lp64", "ilp64"])
def test_library_exports_every_declared_symbol(model):
    from stda_amd.core import LIB_PATH, LIB_PATH_ILP64, load_library
    path = LIB_PATH if model == "lp64" else LIB_PATH_ILP64
    assert os.path.exists(path), "build the libraries first (python -m stda_amd.build)"
    L = load_library(path)
    for name in declared_functions():
        assert hasattr(L, name), name
    # the index type of iaoat / iconf / the -dual atom list
    assert L.stda_hip_index_size() == (4 if model == "lp64" else 8)


def test_error_strings_without_device():
    from stda_amd.core import load_library
    L = load_library()
    assert L.stda_hip_error_string(2).decode().startswith("no CSF")
    assert L.stda_hip_error_string(7).decode() == "matrix power impossible"


def test_create_rejects_null():
    from stda_amd.core import load_library
    L = load_library()
    assert L.stda_hip_create(None, 0) == 1
