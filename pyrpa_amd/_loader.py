"""Loads the in-tree HIP library libgmod.so (fails loudly when it is missing).

Mirrors libs/flibs/_loader.py:1-17 of the reference (fixed library name,
byref helpers).  There is no CPU fallback: every hot-path call goes to the
HIP library."""
import ctypes
import os
from ctypes import byref, c_double, c_int64

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("GMOD_LIB", os.path.join(_HERE, "libgmod.so"))
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libgmod.so not found at {LIB_PATH}: run `python -m pyrpa_amd.build` "
                               "(the HIP library is required; there is no CPU fallback)")
        _lib = ctypes.CDLL(LIB_PATH)
    return _lib


def i64(value):
    """byref-wrapped 64-bit integer scalar (Fortran ABI)."""
    return byref(c_int64(int(value)))


def dbl(value):
    """byref-wrapped double scalar (Fortran ABI)."""
    return byref(c_double(float(value)))


def check():
    """Raise if the library recorded an error in the last call(s)."""
    buf = ctypes.create_string_buffer(1024)
    n = lib().gmod_last_error(buf, 1024)
    if n:
        raise RuntimeError("libgmod: " + buf.value.decode(errors="replace"))
