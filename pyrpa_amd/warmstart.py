"""Warm starts for the FLEX fixed-point loop (mkself with sw_in=True).

The reference seeds mkself from sigma.bin (libs/src/fself.f90:325-340,
742-759) and re-grids a stored seed in frequency in Python
(libs/plibs/_calc.py:52-182).  Here a seed is handed over in memory
(gmod_set_sigma_seed); the plain loop still runs to the same convergence
criterion, so a seed only changes where the iteration starts.

  coarse : a FLEX solve of the same lattice on a coarser k-mesh and fewer
           Matsubara frequencies (same T), re-gridded on the device
           (gmod_regrid_sigma: Fourier interpolation in k, the shared
           Matsubara frequencies kept, c0 + c1/(i w) tail beyond).
"""
from __future__ import annotations

import time

from . import flibs as F
from . import lattice as L


def coarse_mesh(st, factor=2, nw_coarse=None):
    """(Nx, Ny, Nz, Nw) of the coarse solve: every k dimension > 1 divided by
    `factor` (kept >= 4); the Matsubara count of st unless given.  The
    sharp-cutoff self-energy depends on the cutoff at all but the lowest
    frequencies (the circular convolution wraps around), so a seed from fewer
    frequencies is far from the fine solution; a coarser k-mesh at the same
    cutoff lands within ~1e-2 of it."""
    def cut(n):
        return max(4, n // factor) if n > 1 else 1
    nw = nw_coarse if nw_coarse is not None else st['Nw']
    return cut(st['Nx']), cut(st['Ny']), cut(st['Nz']), nw


def coarse_seed(st, factor=2, nw_coarse=None, eps=1e-3, precision=1, scf_loop=60, sw_tail=False):
    """Solve FLEX on the coarse mesh and re-grid Sigma onto st's mesh.
    Returns ((sigma, mu, mu_old), info).  The seed's distance from the fine
    solution is the coarse mesh's discretisation error (~1e-2 at factor 2),
    so the coarse solve stops at eps = 1e-3, an order below it: tighter
    tolerances cost coarse iterations without moving the seed."""
    Nxc, Nyc, Nzc, Nwc = coarse_mesh(st, factor, nw_coarse)
    t0 = time.perf_counter()
    sc = L.build(st['rvec'], st['ham_r'], Nxc, Nyc, Nzc, Nwc, st['temp'], st['U'], st['J'], st['fill'], mu=st['mu'])
    old = F.get_precision()
    try:
        F.set_precision(precision)
        sig_c, mu_c = L.flex(sc, eps=eps, scf_loop=scf_loop, sw_rescale=True, sw_tail=sw_tail)
        stats = F.flex_stats()
    finally:
        F.set_precision(old)
    sig_f = F.regrid_sigma(sig_c, sc['kmap'], sc['invk'], (Nxc, Nyc, Nzc, Nwc), st['kmap'], st['invk'],
                           (st['Nx'], st['Ny'], st['Nz'], st['Nw']), st['temp'])
    info = {"kind": f"coarse FLEX {Nxc}x{Nyc}x{Nzc} x {Nwc} Matsubara (tol {eps:g}"
                    f"{', tail-corrected chi0' if sw_tail else ''}), re-gridded",
            "coarse_iterations": int(stats['iterations']), "seconds": time.perf_counter() - t0}
    return (sig_f, mu_c, mu_c), info


def warm_start(st, kind="coarse", **kw):
    """Seed (sigma, mu, mu_old) for mkself(sw_in=True) on workload `st`, and a
    description.  kind: "coarse" or "none" (returns (None, info))."""
    if kind == "none":
        return None, {"kind": "none (cold start)"}
    if kind == "coarse":
        return coarse_seed(st, **kw)
    raise ValueError(kind)
