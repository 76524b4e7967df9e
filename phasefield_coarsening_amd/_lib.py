"""Locating and loading the package's native libraries.

The libraries are built in-tree (``__graft_entry__.build()``) and loaded from
this directory by absolute path, so a process always runs the files built
here.  A missing library raises: there is no fallback implementation.
"""
import ctypes
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)

_cache = {}


class NativeLibraryMissing(RuntimeError):
    pass


def lib_path(name):
    return os.path.join(PKG_DIR, name)


def load(name):
    """Load ``name`` (e.g. ``libppf.so``) from the package directory once."""
    if name in _cache:
        return _cache[name]
    path = lib_path(name)
    if not os.path.exists(path):
        raise NativeLibraryMissing(
            f"{path} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
            "from the repository root")
    lib = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
    _cache[name] = lib
    return lib
