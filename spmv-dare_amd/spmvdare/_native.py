"""Loading of the in-tree native libraries (ctypes, no PyTorch).

The product path has no CPU fallback: if a library is missing the import of
the module that needs it raises, naming the build command.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}


class NativeMissing(ImportError):
    pass


def load(name):
    """Load libspmv<name>.so from the package directory (cached)."""
    if name in _LIBS:
        return _LIBS[name]
    path = os.path.join(_HERE, "lib%s.so" % name)
    if not os.path.exists(path):
        raise NativeMissing(
            "%s is not built; run `make -C spmv-dare_amd` (or __graft_entry__.build())" % path)
    lib = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
    _LIBS[name] = lib
    return lib


def lib_path(name):
    return os.path.join(_HERE, "lib%s.so" % name)
