"""Real-frequency susceptibility and Pade continuation: the oracle pinned by
the mathematics, and the host-side entry points (get_qshift_, get_tr_chi,
get_qp_) against it.

  * chi0(q, w + i delta) of the band sum (fchi.f90:98-157) continued to the
    imaginary axis (w -> 0+, delta = nu_m) is the Matsubara bubble, which is
    itself pinned through the FFT chi0 and its convergence orders; at w = 0,
    delta = 0 it is the static bubble, -f' = f (1 - f)/T for equal energies.
  * The Thiele continued fraction (fpade.f90) through 2M points of a
    function with M simple poles is that function: the continuation
    reproduces it anywhere in the plane.
  * get_qshift against a brute-force search."""
import numpy as np
import pytest

import oracle.pyrpa_oracle as O
from tests import models as MD

RUN = pytest.mark.parametrize("run", [pytest.param("cpu"), pytest.param("gpu", marks=pytest.mark.gpu)])


def full_mesh(st):
    """klist of the full Nx x Ny x Nz mesh and its bands (the k list the
    real-frequency routines sum over)."""
    Nx, Ny, Nz = st['Nx'], st['Ny'], st['Nz']
    ix, iy, iz = np.meshgrid(np.arange(Nx), np.arange(Ny), np.arange(Nz), indexing='ij')
    kl = np.stack([ix.ravel() / Nx, iy.ravel() / Ny, iz.ravel() / Nz], axis=1)
    eig, uni = O.get_eig(O.gen_ham(kl, st['ham_r'], st['rvec']))
    return kl, eig, uni, O.get_ffermi(eig, st['mu'], st['temp'])


def test_qshift_matches_brute_force():
    st = MD.one_orbital(Nx=6, Ny=5, Nz=2)
    kl, _, _, _ = full_mesh(st)
    for q in ([1 / 6, 2 / 5, 0.5], [0.5, 0.0, 0.0], [5 / 6, 4 / 5, 0.5]):
        qs = O.get_qshift(kl, q)
        for i, k in enumerate(kl):
            t = (k + q) % 1.0
            j = int(np.argmin(np.abs((kl - t + 0.5) % 1.0 - 0.5).sum(axis=1)))
            assert qs[i] == j + 1
    assert (O.get_qshift(kl, [0.01, 0.0, 0.0]) == 1).all()   # off the mesh: index 1


@pytest.mark.parametrize("norb", [1, 2])
def test_band_sum_chi0_continued_to_the_matsubara_axis_is_the_bubble(norb):
    if norb == 1:
        st = MD.one_orbital(Nx=6, Ny=6, Nw=8, temp=0.1, mu=0.2)
    else:
        st = MD.two_orbital(Nx=6, Ny=6, Nw=8, temp=0.1, mu=0.1, V=0.3)
    kl, eig, uni, ff = full_mesh(st)
    ml = [0, 1, 3]
    ex = MD.bubble_exact(st, ml)
    nu = [2 * np.pi * m * st['temp'] for m in ml]
    for iq, q in enumerate(st['klist']):
        qs = O.get_qshift(kl, q)
        for jm, m in enumerate(ml):
            if m == 0:
                c = O._calc_chi(uni, eig, ff, st['olist'], st['temp'], qs, 0.0, 0.0, 0.0)
            else:   # w -> 0+ (the w = 0 branch is the static limit), no weight cut-off
                c = O._calc_chi(uni, eig, ff, st['olist'], st['temp'], qs, 1e-12 * 10, nu[jm], 0.0)
            assert np.abs(c.T - ex[:, :, jm, iq]).max() < 1e-9 * np.abs(ex).max(), (iq, m)


def test_pade_reproduces_a_pole_sum():
    rng = np.random.default_rng(3)
    M = 5
    poles = rng.uniform(-2.0, 2.0, M) - 1j * rng.uniform(0.0, 0.2, M)
    w = rng.uniform(0.2, 1.0, M)

    def R(z):
        return (w[None, :] / (np.asarray(z)[:, None] - poles[None, :])).sum(axis=1)
    T = 0.05
    iw = 1j * (2 * np.arange(2 * M) + 1) * np.pi * T
    wl = np.linspace(-3, 3, 61) + 0.05j
    a = O.get_a(R(iw), iw)
    Q, P = O.get_QP(a, iw, wl)
    # the continuation is exact in exact arithmetic; in FP64 it loses digits
    # near poles 0.02 from the real axis (about 1e-6 of the peak here), while
    # a slip in the recursion gives errors of order one
    assert np.abs(P / Q - R(wl)).max() < 1e-5 * np.abs(R(wl)).max()
    # the continued fraction interpolates the data
    Q, P = O.get_QP(a, iw, iw)
    assert np.abs(P / Q - R(iw)).max() < 1e-10 * np.abs(R(iw)).max()
    # traced over orbital-diagonal blocks
    A = np.zeros((2, 2, len(iw), 3), dtype=complex)
    A[0, 0, :, :] = R(iw)[:, None]
    A[1, 1, :, :] = 2 * R(iw)[:, None]
    A[0, 1, :, :] = 7.0
    B = O.pade_with_trace(A, iw, wl)
    assert np.abs(B - 3 * R(wl)[None, :]).max() < 3e-5 * np.abs(R(wl)).max()


@RUN
def test_host_entry_points_match_oracle(run):
    import pyrpa_amd.flibs as F
    st = MD.two_orbital(Nx=6, Ny=5, Nz=2)
    kl, _, _, _ = full_mesh(st)
    for q in ([1 / 6, 2 / 5, 0.5], [0.0, 0.0, 0.0], [0.01, 0.0, 0.0]):
        assert np.array_equal(F.get_qshift(kl, np.array(q)), O.get_qshift(kl, q))
    rng = np.random.default_rng(8)
    ol, _ = O.get_chi_orb_list(3, [3])
    chis = rng.standard_normal((5, 9, 9)) + 1j * rng.standard_normal((5, 9, 9))
    chi0 = rng.standard_normal((5, 9, 9)) + 1j * rng.standard_normal((5, 9, 9))
    for x, y in zip(F.get_tr_chi(chis, chi0, ol), O.get_tr_chi(chis, chi0, ol)):
        assert np.array_equal(x, y) or np.abs(x - y).max() < 1e-14
    a = rng.standard_normal(12) + 1j * rng.standard_normal(12)
    xl = 1j * np.arange(12)
    wl = np.linspace(-1, 1, 7) + 0.1j
    for x, y in zip(F.get_QP(a, xl, wl), O.get_QP(a, xl, wl)):
        assert np.abs(x - y).max() < 1e-12 * np.abs(y).max()
