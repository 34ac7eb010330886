"""Pins of the oracle's nonlinear FLEX-Eliashberg restatement (ffeliash.f90)
by closed forms, reductions to already-pinned functions and the reference's
own test (tests/test_rpa_flex.py:208-226).  CPU only."""
import numpy as np
import pytest

import oracle.pyrpa_oracle as O
import tests.models as MD


def _bcs_one_orbital(st, D):
    """Exact Nambu pair of a one-orbital band for a real gap D(k, w) (Fortran
    order (Nk, Nw)): G = -(i w + xi) / den, F = -D / den, den = w^2 + xi^2 + D^2."""
    Nw = D.shape[1]
    xi = np.asarray(st['eig'])[:, 0] - st['mu']
    wl = (2 * np.arange(1, Nw + 1) - 1) * np.pi * st['temp']
    den = wl[None, :] ** 2 + xi[:, None] ** 2 + np.abs(D) ** 2
    return -(1j * wl[None, :] + xi[:, None]) / den, -D / den


def _py(a2):
    """(Nk, Nw) Fortran-order one-orbital array -> Python [1, 1, Nw, Nk]."""
    return np.ascontiguousarray(a2[:, :, None, None].T)


def _dwave(st, amp=0.2):
    k = np.asarray(st['klist'])
    Nw = st['Nw']
    d = amp * (np.cos(2 * np.pi * k[:, 0]) - np.cos(2 * np.pi * k[:, 1]))
    return np.repeat(d[:, None], Nw, axis=1).astype(complex)


def test_chi0_conv_ff_one_orbital_equals_direct_sum():
    """For one orbital with a real, k-even singlet gap the legs F^+14, F23 are
    the exact F on the full grid, so chi^FF = -(T/Nkall) sum F(k+q) F(k) is the
    direct O(Nk^2) sum of get_chi0_sum applied to F (the 1e-9 threshold aside)."""
    st = MD.one_orbital(Nx=6, Ny=6, Nw=8, temp=0.07, mu=-0.2, tx=1.0, ty=1.0)
    _, F = _bcs_one_orbital(st, _dwave(st))
    Fp = _py(F)
    ff = O.get_chi0_conv_ff(Fp, st['kmap'], st['invk'], st['olist'], st['plist'], st['temp'], 6, 6, 1)
    ref = O.get_chi0_sum(Fp, st['klist'], st['kmap'], st['invk'], st['olist'], st['temp'], 6, 6, 1)
    assert np.abs(ff - ref).max() < 2e-9
    assert np.abs(ref).max() > 1e-3


def test_chi0sc_with_zero_f_is_the_normal_bubble():
    st = MD.two_orbital(Nx=4, Ny=4, Nw=6, temp=0.1)
    z = np.zeros_like(st['Gk'])
    cs, cc = O.get_chi0sc(st['Gk'], z, st['kmap'], st['invk'], st['olist'], st['plist'], st['temp'], 4, 4, 1)
    gg = O.get_chi0_conv(st['Gk'], st['kmap'], st['invk'], st['olist'], st['temp'], 4, 4, 1)
    assert np.array_equal(cs, gg) and np.array_equal(cc, gg)


@pytest.mark.parametrize("singlet", [True, False])
def test_sc_vertex_reduces_to_the_normal_vertices(singlet):
    """chi0_S = chi0_C = chi0: V_sigma is the FLEX vertex (fUJ.f90:318-365)
    and V_Delta the pairing vertex (fUJ.f90:253-316)."""
    st = MD.two_orbital(Nx=4, Ny=4, Nw=4, temp=0.1, U=0.8, J=0.1)
    chi = O.get_chi0_conv(st['Gk'], st['kmap'], st['invk'], st['olist'], st['temp'], 4, 4, 1)
    Vs, Vd = O.mkV_flex_nosoc(chi, chi, st['Smat'], st['Cmat'], singlet)
    assert np.abs(Vs - O.get_Vsigma_nosoc_flex(chi, st['Smat'], st['Cmat'])).max() < 1e-12
    assert np.abs(Vd - O.get_V_delta(chi, st['Smat'], st['Cmat'], singlet)).max() < 1e-12


def test_sc_vertex_one_orbital_closed_form():
    """One orbital, S = C = U: chi_s = U a/(1 - U a), chi_c = U b/(1 + U b) with
    a = chi0_S, b = chi0_C distinct, so a swapped channel fails."""
    U, a, b = 1.7, 0.31 + 0.02j, 0.12 - 0.01j
    S = np.array([[U]])
    C = np.array([[U]])
    cS = np.full((1, 1, 1, 1), a)
    cC = np.full((1, 1, 1, 1), b)
    Vs, Vd = O.mkV_flex_nosoc(cS, cC, S, C, True)
    _, Vt = O.mkV_flex_nosoc(cS, cC, S, C, False)
    xs, xc = U * a / (1 - U * a), U * b / (1 + U * b)
    vs = U + 1.5 * U * xs + 0.5 * U * xc - 0.5 * U * (U * a + U * b)
    assert abs(Vs[0, 0, 0, 0] - vs) < 1e-13
    assert abs(Vd[0, 0, 0, 0] - (U + 1.5 * U * xs - 0.5 * U * xc)) < 1e-13
    assert abs(Vt[0, 0, 0, 0] - (-0.5 * U * xs - 0.5 * U * xc)) < 1e-13


def test_sc_dyson_updates_hold_the_bcs_pair():
    """The exact one-orbital Nambu pair (G, F) of a real gap is a fixed point of
    both updates: G0 (1 + D F*) = G and -G0 D G* = F (ffeliash.f90:315-384)."""
    st = MD.one_orbital(Nx=6, Ny=6, Nw=8, temp=0.07, mu=-0.2, tx=1.0, ty=1.0)
    D = _dwave(st, 0.3) * (1.0 + 0.1 * np.arange(st['Nw']))[None, :]
    G, F = _bcs_one_orbital(st, D)
    G0 = st['Gk']
    Gn = O.mkgreliah_trs_nsoc(G0, _py(F), _py(D))
    Fn = O.mkfkeliash_trs_nsoc(G0, _py(G), _py(D))
    assert np.abs(Gn - _py(G)).max() < 1e-13 * np.abs(G).max() + 1e-15
    assert np.abs(Fn - _py(F)).max() < 1e-13


def test_sc_dyson_band_rotated_two_orbital():
    """Two bands, H = U diag(xi) U^+, D = U diag(d) U^+ (one real gap per band):
    the updates give U diag(G_n) U^+ and U diag(F_n) U^+ of the per-band pairs;
    F's Hermitian projection (diagonal real, upper = conj lower) leaves a
    Hermitian F as it is."""
    st = MD.two_orbital(Nx=4, Ny=4, Nw=6, temp=0.08, mu=0.1)
    eig = np.asarray(st['eig'])            # [Nk, No]
    uni = np.asarray(st['uni'])            # Python [Nk, No(band), No(orb)] of Fortran uni(orb, band, k)
    Uf = uni.T                             # (orb, band, k)
    Nk, No = eig.shape
    Nw = st['Nw']
    wl = (2 * np.arange(1, Nw + 1) - 1) * np.pi * st['temp']
    dband = np.stack([0.15 * np.ones(Nk), -0.1 * np.ones(Nk)], axis=1)  # per band, real
    xi = eig - st['mu']
    den = wl[None, :, None] ** 2 + xi[:, None, :] ** 2 + dband[:, None, :] ** 2
    gb = -(1j * wl[None, :, None] + xi[:, None, :]) / den
    fb = -dband[:, None, :] / den
    def rot(diag):  # (Nk, Nw, No) band diagonal -> Fortran (Nk, Nw, No, No) orbital matrices
        U = np.transpose(Uf, (2, 0, 1))[:, None]                   # (Nk, 1, orb, band)
        return (U * diag[:, :, None, :]) @ np.conj(np.swapaxes(U, -1, -2))
    Dm = rot(np.repeat(dband[:, None, :], Nw, axis=1).astype(complex))
    Gm, Fm = rot(gb), rot(fb)
    P = lambda X: np.ascontiguousarray(X.T)
    Gn = O.mkgreliah_trs_nsoc(st['Gk'], P(Fm), P(Dm))
    Fn = O.mkfkeliash_trs_nsoc(st['Gk'], P(Gm), P(Dm))
    assert np.abs(Gn - P(Gm)).max() < 1e-12
    assert np.abs(Fn - P(Fm)).max() < 1e-12


def test_zero_seed_and_zero_interaction_stays_zero():
    """tests/test_rpa_flex.py:208-226."""
    st = MD.one_orbital(Nx=4, Ny=4, Nw=3, temp=0.1)
    d0 = np.zeros((1, 1, st['Nw'], len(st['klist'])), dtype=np.complex128)
    D, sig, _ = O.eliashberg(d0, st['Gk'], st['hamk'], st['Smat'], st['Cmat'], st['olist'], st['plist'], st['kmap'],
                             st['invk'], 0.0, st['temp'], 1e-6, 4, 4, 1, 2, 0, sw_sigma=False, sw_Vconst=True,
                             m_diis=1, gap_min=0.0, sw_amp_newton=False)
    assert D.shape == d0.shape and sig.shape == d0.shape
    assert np.allclose(D, 0.0) and np.allclose(sig, 0.0)


def _below_tc(U=2.92):
    st = MD.one_orbital(Nx=8, Ny=8, Nw=16, temp=0.05, mu=-0.3, tx=1.0, ty=1.0, U=U)
    chi, _ = O.get_chi0(st['Smat'], st['Cmat'], st['Gk'], st['olist'], st['kmap'], st['invk'], st['temp'], 8, 8, 1)
    init = O.gap_symms(st['klist'], 1, 1)
    d, lam = O.linearized_eliashberg(chi.copy(), st['Gk'], st['uni'], init, st['Smat'], st['Cmat'], st['olist'],
                                     st['plist'], st['kmap'], st['invk'], 8, 8, 1, st['temp'], 1, eps=1e-10,
                                     itemax=3000, arnoldi_m=0)
    return st, chi, d * (1.764 * st['temp'] / np.abs(d).max()), lam


@pytest.mark.parametrize("mode", [dict(sw_amp_newton=True, m_diis=5), dict(sw_amp_newton=False, m_diis=5)])
def test_converged_gap_solves_the_bcs_gap_equation(mode):
    """Below Tc (lambda = 1.23) the loop with the constant normal-state vertex
    converges to a gap D with D = (T/Nkall) K[F_BCS(D)], F_BCS the exact
    one-orbital Nambu F of D, K the pinned gap convolution with the pinned
    pairing vertex of the normal chi0; the final G is the exact Nambu G."""
    st, chi, d0, lam = _below_tc()
    assert lam > 1.1
    info = {}
    D, sig, G = O.eliashberg(d0.copy(), st['Gk'], st['hamk'], st['Smat'], st['Cmat'], st['olist'], st['plist'],
                             st['kmap'], st['invk'], st['mu'], st['temp'], 1e-12, 8, 8, 1, 600, 1, sw_sigma=False,
                             sw_Vconst=True, gap_min=1e-2, info=info, **mode)
    assert info['status'] == 'converged'
    Dr = D.T[:, :, 0, 0]
    assert np.abs(Dr.imag).max() == 0.0 and np.abs(Dr).max() > 0.1
    Gx, Fx = _bcs_one_orbital(st, Dr)
    Vd = O.get_V_delta(chi, st['Smat'], st['Cmat'], True)
    nd = O.mkdelta_nsoc(_py(Fx), Vd, st['Smat'], st['Cmat'], st['kmap'], st['invk'], st['plist'], st['olist'],
                        8, 8, 1, True) * (st['temp'] / len(st['kmap']))
    assert np.abs(nd[:, :, 0, 0] - Dr).max() < 1e-10 * np.abs(Dr).max()
    assert np.abs(G.T[:, :, 0, 0] - Gx).max() < 1e-10 * np.abs(Gx).max()
    assert np.allclose(sig, 0.0)


def test_gap_collapses_above_tc():
    st, chi, d0, lam = _below_tc(U=2.6)
    assert lam < 1.0
    info = {}
    D, _, _ = O.eliashberg(d0.copy(), st['Gk'], st['hamk'], st['Smat'], st['Cmat'], st['olist'], st['plist'],
                           st['kmap'], st['invk'], st['mu'], st['temp'], 1e-10, 8, 8, 1, 300, 1, sw_sigma=False,
                           sw_Vconst=True, gap_min=1e-2, info=info, sw_amp_newton=False, m_diis=5)
    assert info['status'] == 'collapsed'
    assert np.abs(D).max() < 1e-2 * np.abs(d0).max()
