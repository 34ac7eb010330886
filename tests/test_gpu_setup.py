"""H(k) (fmod.f90:162-207) and the band eigenpairs (fmod.f90:209-248) formed
on the device, against the oracle (numpy: the reference's formula for H(k),
LAPACK's Hermitian eigen-solver): H(k) to rounding; eigenvalues to 1e-12;
eigenvectors through phase-free identities (U^H U = I, U diag(e) U^H = H)
and, for non-degenerate bands, the band projectors U_an conj(U_bn)."""
import time

import numpy as np
import pytest

import oracle.pyrpa_oracle as O
from tests import models as MD

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def F():
    import pyrpa_amd.flibs as F
    return F


CASES = {
    "1orb_8x8": lambda: MD.one_orbital(Nx=8, Ny=8, Nw=2, temp=0.1),
    "2orb_6x6": lambda: MD.two_orbital(Nx=6, Ny=6, Nw=2, temp=0.1),
    "3orb_4x4x2_complex": lambda: MD.random_model(3, 4, 4, 2, 2, 0.1, 0.0, 0.0, 0.0, seed=11),
    "4orb_4x4x2": lambda: MD.random_model(4, 4, 4, 2, 2, 0.1, 0.0, 1.0, 0.2, seed=1),
    "5orb_4x4": lambda: MD.random_model(5, 4, 4, 1, 2, 0.1, 0.0, 0.0, 0.0, seed=3),
    "8orb_2x2x2_complex": lambda: MD.random_model(8, 2, 2, 2, 2, 0.1, 0.0, 0.0, 0.0, seed=5),
}


def _check_eig(hk, e, u, e_ref, u_ref):
    assert np.abs(e - e_ref).max() < 1e-12
    U = np.transpose(u, (0, 2, 1))          # Python [k, row, column] of Fortran uni(row, column, k)
    H = np.transpose(hk, (0, 2, 1))
    No = U.shape[1]
    assert np.abs(np.einsum('kmn,kml->knl', np.conj(U), U) - np.eye(No)).max() < 1e-12
    assert np.abs(np.einsum('kmn,kn,kln->kml', U, e, np.conj(U)) - H).max() < 1e-12
    Ur = np.transpose(u_ref, (0, 2, 1))
    gap = np.diff(e_ref, axis=1)
    for k in range(U.shape[0]):
        for n in range(No):
            if (n > 0 and gap[k, n - 1] < 1e-6) or (n < No - 1 and gap[k, n] < 1e-6):
                continue  # degenerate band: its projector is not unique
            P = np.outer(U[k, :, n], np.conj(U[k, :, n]))
            Pr = np.outer(Ur[k, :, n], np.conj(Ur[k, :, n]))
            assert np.abs(P - Pr).max() < 1e-10


@pytest.mark.parametrize("name", list(CASES))
def test_gen_ham_and_bands_match_oracle(F, name):
    st = CASES[name]()
    hk = F.gen_ham(st['klist'], st['ham_r'], st['rvec'])
    assert np.abs(hk - st['hamk']).max() < 1e-14
    e, u = F.get_eig(hk)
    _check_eig(hk, e, u, st['eig'], st['uni'])


def test_degenerate_bands():
    """Orbital-diagonal H with equal on-site terms: exact degeneracies, eigenvalues
    still ascending and the eigenvectors an orthonormal basis of each eigenspace."""
    import pyrpa_amd.flibs as F
    rng = np.random.default_rng(2)
    No, Nk = 4, 50
    d = np.repeat(rng.uniform(-1, 1, (Nk, 2)), 2, axis=1)          # pairs of equal levels
    Q = np.linalg.qr(rng.standard_normal((Nk, No, No)) + 1j * rng.standard_normal((Nk, No, No)))[0]
    H = np.einsum('kmn,kn,kln->kml', Q, d, np.conj(Q))              # [k, row, col]
    hk = np.ascontiguousarray(np.transpose(H, (0, 2, 1)))
    e, u = F.get_eig(hk)
    assert np.abs(e - np.sort(d, axis=1)).max() < 1e-12
    U = np.transpose(u, (0, 2, 1))
    assert np.abs(np.einsum('kmn,kml->knl', np.conj(U), U) - np.eye(No)).max() < 1e-12
    assert np.abs(np.einsum('kmn,kn,kln->kml', U, e, np.conj(U)) - H).max() < 1e-12


def test_initial_mu_mesh_on_the_device(F):
    """The 40^3 mesh of the initial chemical potential (libs/plibs/_calc.py:298-309)
    for the 4-orbital 3D lattice: 64000 k points through the device."""
    from pyrpa_amd import lattice as L
    st = MD.random_model(4, 2, 2, 2, 2, 0.1, 0.0, 0.0, 0.0, seed=9, real=True)
    g = np.arange(40) / 40
    kx, ky, kz = np.meshgrid(g, g, g, indexing='ij')
    kl = np.stack([kx.ravel(), ky.ravel(), kz.ravel()], axis=1)
    t0 = time.perf_counter()
    hk = F.gen_ham(kl, st['ham_r'], st['rvec'])
    e, u = F.get_eig(hk)
    print(f"\n[setup] 64000 k points: H(k) and bands {time.perf_counter() - t0:.3f} s")
    assert np.abs(hk - O.gen_ham(kl, st['ham_r'], st['rvec'])).max() < 1e-13
    er, ur = O.get_eig(hk)
    assert np.abs(e - er).max() < 1e-12
