"""pyrpa_amd: MI355X-native chi0/RPA, FLEX and linearized-Eliashberg engine.

``pyrpa_amd.flibs`` mirrors the reference's ``libs.flibs`` ctypes layer over
the HIP library ``libgmod.so`` (built in-tree by ``python -m pyrpa_amd.build``).
``pyrpa_amd.lattice`` holds the host-side model builders and helpers.
"""
