"""Locate and load the package's in-tree shared libraries."""
from __future__ import annotations

import ctypes as C
import os

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")

_cache: dict[str, C.CDLL] = {}


class LibraryMissing(RuntimeError):
    pass


def lib_path(name: str) -> str:
    return os.path.join(LIB_DIR, name)


def load_host_lib(name: str) -> C.CDLL:
    """Load lib/<name>; raise LibraryMissing (never fall back) if it is absent."""
    if name in _cache:
        return _cache[name]
    path = lib_path(name)
    if not os.path.exists(path):
        raise LibraryMissing(f"{path} is not built: run `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = C.CDLL(path)
    _cache[name] = lib
    return lib
