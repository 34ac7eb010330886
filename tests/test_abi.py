"""Checks of the C-ABI library: it loads, exports every symbol declared in
include/gmod.h, and its host-side setup entry points match the oracle.

The setup entry points run on the host (no GPU call), so their tests run in
the CPU suite and again, through the RUN parameter's gpu-marked case, in the
GPU suite on the machine the library is measured on."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle.pyrpa_oracle as O
from tests import models as MD

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "gmod.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:void|int|int64_t|double)\s*\*?\s+(\w+)\s*\(", txt, flags=re.M)))


# every host-side setup test runs once in each suite
RUN = pytest.mark.parametrize("run", [pytest.param("cpu"), pytest.param("gpu", marks=pytest.mark.gpu)])


def test_header_declares_the_hot_path_entry_points():
    syms = declared_symbols()
    for name in ["mkself", "get_chi0", "get_chi0_tail", "get_chi0_conv_", "get_vsigma_flex_nosoc_",
                 "get_chis", "get_chis_chic", "get_eig_or_tr_chi", "gen_green0_", "gen_green_inv_", "getinv_",
                 "lin_eliash", "mkfk_trs_nsoc_", "get_initial_delta_", "conv_delta_orb_to_band", "remap_delta_",
                 "get_scmat", "get_scmat_orb", "generate_irr_kpoint_inv", "gen_ham", "get_eig", "get_ffermi",
                 "openmp_params", "gmod_last_error", "gmod_get_chis_chic_all", "gmod_nn_create",
                 "gmod_nn_train_epoch", "mkself_soc", "get_chi0_soc", "get_chis_chic_soc", "get_vmat_soc",
                 "get_vmat_soc_orb", "get_chi_irr", "chiq_map", "pade_with_trace", "gmod_wannier_r",
                 "eliashberg", "get_chi0sc_", "mkv_flex_nosoc_"]:
        assert name in syms


def test_library_loads_and_exports_every_declared_symbol():
    from pyrpa_amd import _loader
    lib = _loader.lib()
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


@RUN
@pytest.mark.parametrize("mesh", [(2, 2, 1), (4, 4, 1), (8, 8, 1), (4, 4, 4), (3, 3, 1), (4, 3, 1),
                                  (6, 4, 2), (5, 5, 3), (4, 4, 3)])
def test_irreducible_mesh_matches_oracle(mesh, run):
    import pyrpa_amd.flibs as F
    a = F.gen_irr_k_TRS(*mesh)
    b = O.gen_irr_k_TRS(*mesh)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


@RUN
def test_irreducible_mesh_full_size_consistent(run):
    """32x32x16 (the 4-orbital 3D mesh): pairing invariants at full size."""
    import pyrpa_amd.flibs as F
    klist, kmap, invk = F.gen_irr_k_TRS(32, 32, 16)
    assert len(klist) == 8196
    N = np.array([32, 32, 16])
    kc = np.rint(klist * N).astype(int)
    irr = invk[:, 0] - 1
    d = invk[:, 1] == 0
    assert np.array_equal(kc[irr[d]], kmap[d])
    assert np.array_equal((-kc[irr[~d]]) % N, kmap[~d])
    assert np.array_equal(kmap[invk[:, 2] - 1], (-kmap) % N)
    assert np.bincount(irr, minlength=len(klist)).sum() == 32 * 32 * 16


@RUN
def test_setup_host_entry_points_match_oracle(run):
    import pyrpa_amd.flibs as F
    st = MD.random_model(4, 4, 4, 2, 4, 0.1, 0.0, 1.0, 0.2, seed=1)
    e = st['eig']  # H(k) and the band eigenpairs are formed on the device (tests/test_gpu_setup.py)
    ol, site = O.get_chi_orb_list(2, [2])
    for a, b in zip(F.gen_SCmatrix(ol, site, 1.0, 0.2), O.get_scmat(ol, site, 1.0, 0.2)):
        assert np.array_equal(a, b)
    Um, Jm = np.array([[1.0, 0.8], [0.8, 1.2]]), np.array([[0.0, 0.2], [0.2, 0.0]])
    for a, b in zip(F.gen_SCmatrix_orb(ol, site, Um, Jm), O.get_scmat_orb(ol, site, Um, Jm)):
        assert np.allclose(a, b)
    assert np.abs(F.get_ffermi(e, 0.1, 0.05) - O.get_ffermi(e, 0.1, 0.05)).max() < 1e-15
    w = F.gen_kpoint_weight(st['invk'], len(st['klist']))
    assert np.array_equal(w, O.kweight(st['invk'], len(st['klist'])))
    # the spinful vertex (its spin-degenerate channels are pinned in test_soc_oracle.py);
    # the orbital-dependent form with uniform matrices is the uniform one
    for Ns in (2, 4, 6):
        slist, invs = O.soc_spin_lists(Ns)
        ol, site = O.get_chi_orb_list(Ns, [Ns])
        V = F.gen_Vmatrix(ol, slist, site, invs, 1.1, 0.25)
        assert np.array_equal(V, O.get_vmat_soc(ol, slist, site, invs, 1.1, 0.25))
        Um = np.full((Ns, Ns), 1.1 - 2 * 0.25)
        for a in range(Ns):
            Um[a, a] = Um[a, invs[a] - 1] = 1.1
        Vo = F.gen_Vmatrix_orb(ol, slist, site, invs, Um, np.full((Ns, Ns), 0.25))
        assert np.array_equal(Vo, V)


@RUN
def test_gap_output_formatting_matches_oracle(run):
    """conv_delta_orb_to_band / remap_gap are host-side output formatting."""
    import pyrpa_amd.flibs as F
    st = MD.random_model(3, 4, 4, 1, 3, 0.1, 0.0, 1.0, 0.2, seed=2)
    rng = np.random.default_rng(3)
    Nk = len(st['klist'])
    delta = rng.standard_normal((3, 3, 3, Nk)) + 1j * rng.standard_normal((3, 3, 3, Nk))
    prt = np.array([1.0, -1.0, 1.0])
    for gs in (1, -1):
        a = F.conv_delta_orb_to_band(delta, st['uni'], st['invk'], prt, gs)
        b = O.conv_delta_orb_to_band(delta, st['uni'], st['invk'], prt, gs)
        assert np.abs(a - b).max() < 1e-13
        a = F.remap_gap(delta, prt, st['invk'], gs)
        b = O.remap_gap(delta, prt, st['invk'], gs)
        assert np.abs(a - b).max() < 1e-15
