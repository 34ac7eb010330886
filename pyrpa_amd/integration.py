"""Routing of the reference's ctypes handle to libgmod.so.

The reference's wrappers bind one handle, ``_lib``, created in
``libs/flibs/_loader.py:6-7`` and imported by every wrapper module
(``from ._loader import _lib``, e.g. ``libs/flibs/_selfenergy.py:3``), which
then set ``argtypes`` on each symbol at import time.  ``route`` returns an
object that stands in for that handle: the symbols of the FFT-convolved
many-body path resolve in libgmod.so, every other symbol in the original
library, so the reference's argtypes land on the function that is called.
It is installed by rebinding ``_lib`` right after the original load::

    _lib = np.ctypeslib.load_library("libfmod.so", str(_LIB_DIR))
    from pyrpa_amd.integration import route
    _lib = route(_lib)            # no-op unless PYRPA_GMOD names libgmod.so
"""
import ctypes
import os

# the libfmod.so symbols libgmod.so replaces (include/gmod.h)
HOT = frozenset((
    # FLEX, chi0 and the linearized gap equation
    "mkself", "get_chi0", "get_chi0_tail", "get_chi0_conv_", "get_vsigma_flex_nosoc_", "get_chis", "get_chis_chic",
    "get_eig_or_tr_chi", "gen_green0_", "gen_green_inv_", "getinv_", "lin_eliash", "mkfk_trs_nsoc_",
    "get_initial_delta_", "conv_delta_orb_to_band", "remap_delta_",
    # spinful basis
    "mkself_soc", "get_chi0_soc", "get_chis_chic_soc", "lin_eliash_soc", "mkfk_trs_soc_", "conv_delta_orb_to_band_soc",
    "get_vmat_soc", "get_vmat_soc_orb",
    # real-frequency maps and Pade continuation
    "get_chi_irr", "chiq_map", "get_tr_chi", "get_qshift_", "get_a_", "get_qp_", "pade_with_trace",
    # nonlinear FLEX-Eliashberg
    "eliashberg",
))


class Router:
    """Attribute access resolves hot-path symbols in `gmod`, the rest in `fallback`.
    ctypes caches each function object on its library, so argtypes set through
    the router stay with the function the router returns next time."""

    def __init__(self, gmod, fallback, hot=HOT):
        self._gmod = gmod
        self._fallback = fallback
        self._hot = frozenset(hot)

    def __getattr__(self, name):
        if name.startswith("__"):
            raise AttributeError(name)
        return getattr(self._gmod if name in self._hot else self._fallback, name)


def route(fallback, gmod_path=None):
    """`fallback` itself when no libgmod.so is named (argument or the
    PYRPA_GMOD environment variable), else a Router over both libraries."""
    path = gmod_path or os.environ.get("PYRPA_GMOD")
    if not path:
        return fallback
    return Router(ctypes.CDLL(path), fallback)
