"""Locating and loading the in-tree native libraries.

The HIP engine (`libsilo_amd.so`) and the host-only data generator (`libsilo_datagen.so`) are
built in-tree next to their sources by `__graft_entry__.build()`; nothing is installed into
site-packages, and a missing library is an error, never a silent fallback.
"""
from __future__ import annotations

import ctypes
import os

PACKAGE_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(PACKAGE_ROOT, "csrc")
ENGINE_LIB = os.path.join(CSRC, "libsilo_amd.so")
DATAGEN_LIB = os.path.join(CSRC, "libsilo_datagen.so")

_cache: dict[str, ctypes.CDLL] = {}


class NativeLibraryMissing(RuntimeError):
    pass


def load(path: str) -> ctypes.CDLL:
    lib = _cache.get(path)
    if lib is None:
        if not os.path.exists(path):
            raise NativeLibraryMissing(
                f"{path} is not built; run __graft_entry__.build() (hipcc for gfx950) first"
            )
        lib = ctypes.CDLL(path)
        _cache[path] = lib
    return lib
