"""MI355X (gfx950) Box2D step: Python loaders for the in-tree libraries.

The product is C/C++: libb2g.so (HIP kernels behind include/b2g.h) and
libbox2d.so (the box2d:: host API plus a C scene runner). Python only loads
them for tests and the benchmark; there is no Python or CPU fallback for the
step — if the libraries are missing, loading fails loudly.
"""
from __future__ import annotations

import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIB_DIR = os.path.join(PKG, "lib")


class LibraryMissing(RuntimeError):
    pass


def _load(name: str) -> ctypes.CDLL:
    path = os.path.join(LIB_DIR, name)
    if not os.path.exists(path):
        raise LibraryMissing(f"{path} is not built; run `python -m box2dpp_amd.build`")
    return ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)


_b2g = None
_host = None


def b2g() -> ctypes.CDLL:
    """The HIP step library (C ABI of include/b2g.h)."""
    global _b2g
    if _b2g is None:
        _b2g = _load("libb2g.so")
    return _b2g


def host() -> ctypes.CDLL:
    """The box2d:: host API library with the b2pp_* scene runner."""
    global _host
    if _host is None:
        b2g()
        _host = _load("libbox2d.so")
        _host.b2pp_world_create.restype = ctypes.c_void_p
        _host.b2pp_world_create.argtypes = [ctypes.c_float, ctypes.c_float]
        _host.b2pp_world_last_error.restype = ctypes.c_char_p
        _host.b2pp_world_last_error.argtypes = [ctypes.c_void_p]
        _host.b2pp_world_device_context.restype = ctypes.c_void_p
        _host.b2pp_world_device_context.argtypes = [ctypes.c_void_p]
    return _host
