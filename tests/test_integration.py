"""The drop-in routing of the reference's ctypes handle (pyrpa_amd.integration):
hot-path symbols resolve in libgmod.so, others in the original library, and
argtypes set through the handle, as the reference's wrapper modules do at
import (libs/flibs/_selfenergy.py:7-37), stay on the function that is then
called.  Host-only: the routed call below is get_scmat, a host entry point."""
import ctypes

import numpy as np

import oracle.pyrpa_oracle as O
from pyrpa_amd import _loader, integration


class _Original:
    """Stands in for libfmod.so's handle: the symbols libgmod.so does not
    replace (here one of the reference's band-structure entry points)."""

    def __init__(self):
        self.get_vlm0 = object()


def test_route_without_libgmod_returns_the_original_handle(monkeypatch):
    monkeypatch.delenv("PYRPA_GMOD", raising=False)
    orig = _Original()
    assert integration.route(orig) is orig


def test_router_resolves_hot_symbols_in_libgmod_and_keeps_argtypes(monkeypatch):
    monkeypatch.setenv("PYRPA_GMOD", _loader.LIB_PATH)
    orig = _Original()
    lib = integration.route(orig)
    assert lib.get_vlm0 is orig.get_vlm0
    gm = ctypes.CDLL(_loader.LIB_PATH)
    for name in integration.HOT:
        assert hasattr(gm, name), name
    # the reference's wrappers set argtypes once, at import, through the handle
    nd = np.ctypeslib.ndpointer
    lib.mkself.argtypes = [nd(dtype=np.complex128)] + [ctypes.c_void_p] * 30
    assert lib.mkself.argtypes is not None and len(lib.mkself.argtypes) == 31
    assert lib.mkself is lib.mkself


def test_a_call_through_the_router_reaches_libgmod(monkeypatch):
    """get_scmat is not on the hot list: route it explicitly to show the call
    path (handle -> CDLL function with the wrapper's argtypes -> result)."""
    gm = ctypes.CDLL(_loader.LIB_PATH)
    lib = integration.Router(gm, _Original(), hot={"get_scmat"})
    nd = np.ctypeslib.ndpointer
    P = ctypes.POINTER
    lib.get_scmat.argtypes = [nd(dtype=np.float64), nd(dtype=np.float64), nd(dtype=np.int64),
                              nd(dtype=np.int64), P(ctypes.c_double), P(ctypes.c_double), P(ctypes.c_int64)]
    lib.get_scmat.restype = None
    ol, site = O.get_chi_orb_list(2, [2])
    S = np.zeros((4, 4))
    C = np.zeros((4, 4))
    # ol is F-contiguous: its memory is the Fortran olist(Nchi,2)
    lib.get_scmat(S, C, ol, site, ctypes.byref(ctypes.c_double(1.0)), ctypes.byref(ctypes.c_double(0.2)),
                  ctypes.byref(ctypes.c_int64(4)))
    Sr, Cr = O.get_scmat(ol, site, 1.0, 0.2)
    assert np.array_equal(S, Sr) and np.array_equal(C, Cr)
