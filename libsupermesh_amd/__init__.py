"""MI355X-native supermeshing engine exposing the libsupermesh C ABI.

Mirrors the reference's Python loader (src/supermesh/__init__.py:8-26):
``get_library()`` / ``get_include()`` / ``load()``.  The loader fails loudly
when the HIP library has not been built -- there is no CPU fallback.
"""
import ctypes
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
_ROOT = os.path.dirname(_PKG)


def get_library():
    """Path of the built shared library (libsupermesh_gpu.so)."""
    path = os.path.join(_PKG, "lib", "libsupermesh_gpu.so")
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build it with `python -m libsupermesh_amd.build` "
            "(the library has no CPU fallback)")
    return path


def get_include():
    """Directory holding libsupermesh-c.h and libsupermesh_gpu.h."""
    return os.path.join(_ROOT, "include")


def load():
    """Load the shared library and return a ctypes handle."""
    return ctypes.CDLL(get_library())
