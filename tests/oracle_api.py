"""Loads the FP64 CPU oracle (oracle/liboracle.so) for tests; builds it when absent."""
import ctypes
import os
import subprocess

from topoptspline3d_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "liboracle.so")
_api = None


def oracle_api():
    global _api
    if _api is None:
        if not os.path.exists(LIB):
            subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
        _api = _abi.bind(ctypes.CDLL(LIB), "orc_")
    return _api
