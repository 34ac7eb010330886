"""The product's host helpers against the oracle and the mathematics:
olist (libs/plibs/_response.py:382-413), the initial-gap form factors
(_response.py:415-445) and the initial chemical potential (_bands.py:50-78,
_calc.py:298-309).  The host-only helpers run in the CPU suite and again,
through the gpu-marked case of RUN, in the GPU suite; the initial chemical
potential needs the bands, which are formed on the device."""
import numpy as np
import pytest

import oracle.pyrpa_oracle as O
from tests import models as MD

RUN = pytest.mark.parametrize("run", [pytest.param("cpu"), pytest.param("gpu", marks=pytest.mark.gpu)])


@RUN
@pytest.mark.parametrize("norb", [1, 2, 3, 4, 5])
def test_chi_orb_list_single_site(norb, run):
    from pyrpa_amd import lattice as L
    a, sa = L.get_chi_orb_list(norb)
    b, sb = O.get_chi_orb_list(norb, [norb])
    assert a.dtype == np.int64 and np.array_equal(a, b) and np.array_equal(sa, sb)
    # the memory is the Fortran olist(Nchi,2): first orbitals, then second orbitals
    assert a.flags.f_contiguous
    mem = np.frombuffer(a.tobytes(order='A'), dtype=np.int64)
    assert np.array_equal(mem[:norb * norb], np.tile(np.arange(1, norb + 1), norb))
    assert np.array_equal(mem[norb * norb:], np.repeat(np.arange(1, norb + 1), norb))


@RUN
def test_chi_orb_list_multi_site(run):
    """Two sites of 1 and 2 orbitals: per-site blocks, site label = first
    orbital of the site, C-contiguous like the reference's list stack."""
    from pyrpa_amd import lattice as L
    ol, site = L.get_chi_orb_list(3, [1, 2])
    assert np.array_equal(ol, [[1, 1], [2, 2], [3, 2], [2, 3], [3, 3]])
    assert np.array_equal(site, [1, 2, 2, 2, 2])
    assert ol.flags.c_contiguous
    with pytest.raises(ValueError):
        L.get_chi_orb_list(4, [1, 2])


@RUN
def test_gap_form_factors(run):
    from pyrpa_amd import lattice as L
    rng = np.random.default_rng(5)
    kl = rng.uniform(0, 1, (40, 3))
    for gs in (0, 1, 2, 3, 4, 5, -1, -2, -3, 7):
        a = L.gap_symms(kl, 3, gs)
        b = O.gap_symms(kl, 3, gs)
        assert a.shape == (3, 40) and np.allclose(a, b, rtol=0, atol=1e-15)
    # d_{x2-y2} is odd and d_xy even under kx <-> ky; p_x + i p_y = p_x + i p_y
    sw = kl[:, [1, 0, 2]]
    assert np.allclose(L.gap_symms(sw, 1, 1), -L.gap_symms(kl, 1, 1))
    assert np.allclose(L.gap_symms(sw, 1, 3), L.gap_symms(kl, 1, 3))
    assert np.allclose(L.gap_symms(kl, 1, -3), L.gap_symms(kl, 1, -1) + 1j * L.gap_symms(kl, 1, -2))
    assert np.allclose(L.gap_symms(kl, 2, 0), 1.0)


@pytest.mark.gpu
def test_initial_mu():
    from pyrpa_amd import lattice as L
    # particle-hole symmetric nearest-neighbour square lattice at half filling: mu = 0
    rvec, ham_r = L.square_lattice(1.0, 0.0)
    assert abs(L.get_mu(ham_r, rvec, 0.05, 0.5, kmesh=16)) < 1e-10
    assert abs(O.get_mu(ham_r, rvec, 0.05, 0.5, kmesh=16)) < 1e-10
    # t' breaks the symmetry: product and oracle agree, and the oracle's mu fills the band
    rvec, ham_r = L.square_lattice(1.0, -0.25)
    a = L.get_mu(ham_r, rvec, 0.02, 0.425, kmesh=24)
    b = O.get_mu(ham_r, rvec, 0.02, 0.425, kmesh=24)
    assert abs(a - b) < 1e-9
    g = np.arange(24) / 24
    kx, ky, kz = np.meshgrid(g, g, g, indexing='ij')
    kl = np.stack([kx.ravel(), ky.ravel(), kz.ravel()], axis=1)
    e, _ = O.get_eig(O.gen_ham(kl, ham_r, rvec))
    assert abs(O.get_ffermi(e, b, 0.02).sum() / len(kl) - 0.425) < 1e-12
    # a filling beyond the band range is clamped to the closer band edge
    assert L.calc_mu(e, len(kl), 1.5, 0.02) == pytest.approx(e.max())
    assert L.calc_mu(e, len(kl), -0.5, 0.02) == pytest.approx(e.min())
    # a multi-orbital lattice
    st = MD.random_model(3, 4, 4, 1, 4, 0.1, 0.0, 0.0, 0.0, seed=3)
    a = L.get_mu(st['ham_r'], st['rvec'], 0.05, 1.1, kmesh=12)
    b = O.get_mu(st['ham_r'], st['rvec'], 0.05, 1.1, kmesh=12)
    assert abs(a - b) < 1e-9
