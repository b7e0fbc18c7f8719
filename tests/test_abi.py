"""CPU checks of the boundary: the HIP library loads, exports every entry point that
include/tos_api.h declares, and the ctypes table matches the header."""
import ctypes
import os
import re

import pytest

from topoptspline3d_amd import _abi
from topoptspline3d_amd.backend import LIB

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "tos_api.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(tos_[a-z_0-9]+)\s*\(", src)))


def test_header_and_table_agree():
    names = declared()
    assert names == sorted(_abi.exported_names("tos_"))


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.skip("libtos.so not built")
    lib = ctypes.CDLL(LIB)
    for n in declared():
        assert hasattr(lib, n), n


def test_oracle_exports_the_same_abi():
    from oracle_api import oracle_api
    api = oracle_api()
    for n in _abi.SIGNATURES:
        assert hasattr(api, n)
