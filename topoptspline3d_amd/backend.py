"""Loads the in-tree gfx950 HIP library (topoptspline3d_amd/libtos.so).

There is no fallback: a missing library is an error, so nothing silently runs
anywhere but the GPU."""
import ctypes
import os

from . import _abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libtos.so")
_api = None


def hip_api():
    global _api
    if _api is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} is missing: build it with `python -m topoptspline3d_amd.build`")
        _api = _abi.bind(ctypes.CDLL(LIB), "tos_", learned=True)
    return _api


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
