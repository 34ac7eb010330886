"""Warm starts for the FLEX fixed-point loop (mkself with sw_in=True).

The reference seeds mkself from sigma.bin (libs/src/fself.f90:325-340,
742-759) and re-grids a stored seed in Python (libs/plibs/_calc.py:52-182).
Here a seed is handed over in memory (gmod_set_sigma_seed); the plain loop
still runs to the same convergence criterion, so a seed only changes where
the iteration starts.
"""
from __future__ import annotations


def warm_start(st):
    """Seed (sigma, mu, mu_old) for mkself(sw_in=True) on workload `st`, and a
    description.  Returns (None, info) for a cold start."""
    return None, {"kind": "none (cold start)"}
