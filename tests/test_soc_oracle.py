"""The oracle's spin-orbit (spinful) path pinned by the mathematics
(fchi0withsoc.f90, fselfwithsoc.f90, fUJ.f90:109-251, 367-399).

  * Spin-degenerate reduction: for H(k) (x) 1_spin the spinful chi0 is the
    spinless chi0 on every equal-spin block and zero elsewhere, and the SOC
    vertex of get_vmat_soc(U, J) gives back the Kanamori charge and spin
    channels: chic = chi_c, 2 chiszz = chi_s, chispm = chi_s.  A sign or index
    slip in the vertex table, the orbital map or the RPA algebra breaks them.
  * The spinful k / -k unfolding equals G computed at -k directly, for a
    model with atomic and Rashba-like spin-orbit terms (time reversal only).
  * chi0 of that model against a direct circular sum over the full grid
    (every (l, m), partner copies and their signs included).
  * The vertex: -V + V (I + chi0 V)^-1 chi0 V = -V (I + chi0 V)^-1.
  * Sigma against a direct circular (q, nu) sum.
  * mkself_soc: the returned Sigma is a fixed point of the loop map to its
    tolerance (the mixing, the mu search and n(mu) are the spinless loop's,
    pinned with it)."""
import numpy as np
import pytest

import oracle.pyrpa_oracle as O
from tests import models as MD


def _args(st):
    return (st['kmap'], st['invk'], st['invs'], st['olist'], st['slist'], st['temp'], st['Nx'], st['Ny'], st['Nz'])


@pytest.fixture(scope="module")
def degenerate():
    base = MD.two_orbital(Nx=4, Ny=4, Nw=8, temp=0.1, mu=0.1, V=0.3, U=1.0, J=0.2)
    return base, MD.spinful(base, U=1.0, J=0.2)


@pytest.fixture(scope="module")
def soc():
    base = MD.two_orbital(Nx=4, Ny=4, Nw=8, temp=0.1, mu=0.1, V=0.3)
    return MD.spinful(base, lam=0.4, alpha=0.15, U=0.8, J=0.15)


def test_spin_degenerate_chi0_is_the_spinless_chi0(degenerate):
    base, st = degenerate
    No = base['Norb']
    chi = O.get_chi0_conv_soc(st['Gk'], *_args(st)).T
    chin = O.get_chi0_conv(base['Gk'], base['kmap'], base['invk'], base['olist'], base['temp'],
                           base['Nx'], base['Ny'], base['Nz']).T
    ol, oln = O.olist_f(st['olist']), O.olist_f(base['olist'])
    idx = {tuple(p): i for i, p in enumerate(oln)}
    for l, (a, b) in enumerate(ol):
        for m, (c, d) in enumerate(ol):
            if a // No == c // No and d // No == b // No:
                ref = chin[:, :, idx[(a % No, b % No)], idx[(c % No, d % No)]]
                assert np.abs(chi[:, :, l, m] - ref).max() < 1e-14
            else:
                assert np.abs(chi[:, :, l, m]).max() == 0.0


def test_soc_vertex_gives_kanamori_channels(degenerate):
    base, st = degenerate
    chi = O.get_chi0_conv_soc(st['Gk'], *_args(st))
    chin = O.get_chi0_conv(base['Gk'], base['kmap'], base['invk'], base['olist'], base['temp'],
                           base['Nx'], base['Ny'], base['Nz'])
    chic, szz, spm = O.get_chis_chic_soc(chi, st['Vmat'], st['olist'], st['slist'], st['invs'])
    chis_n, chic_n = O.get_chis_chic(chin, base['Smat'], base['Cmat'])
    scale = np.abs(chis_n).max()
    assert np.abs(chic - chic_n).max() < 1e-13 * scale
    assert np.abs(2 * szz - chis_n).max() < 1e-13 * scale
    assert np.abs(spm - chis_n).max() < 1e-13 * scale
    # and the Stoner factor of -chi0 V is that of chi0 S
    assert abs(O.ckchi_soc(chi, st['Vmat']) - O.ckchi(chin, base['Smat'], base['Cmat'])) < 1e-12


def test_spinful_unfolding_is_time_reversal(soc):
    st = soc
    full = MD.full_green(st)
    Gf = np.asarray(st['Gk']).T
    sg, _ = O.soc_signs(st['olist'], st['slist'])
    inv = st['invs'] - 1
    Ns = st['Norb']
    assert np.abs(full).max() > 0 and np.abs(full[0, Ns // 2]).max() > 1e-3   # spin-mixing present
    for a in range(Ns):
        for b in range(Ns):
            u = O._full_green_soc(Gf, a, b, sg, inv, st['kmap'], st['invk'], st['Nx'], st['Ny'], st['Nz'])
            assert np.abs(u - full[a, b]).max() < 1e-13


def _chi0_direct(st):
    full = MD.full_green(st)
    ol = O.olist_f(st['olist'])
    Nx, Ny, Nz, M = full.shape[2:]
    Nw = M // 2
    Nkall = Nx * Ny * Nz
    d = np.nonzero(st['invk'][:, 1] == 0)[0]
    out = np.zeros((len(st['klist']), Nw, len(ol), len(ol)), dtype=complex)
    for l, (l1, l2) in enumerate(ol):
        for m, (m1, m2) in enumerate(ol):
            g13, g42 = full[l1, m1], full[m2, l2]
            for i in d:
                q = st['kmap'][i]
                sh = np.roll(g13, shift=(-q[0], -q[1], -q[2]), axis=(0, 1, 2))
                for n in range(Nw):
                    out[st['invk'][i, 0] - 1, n, m, l] = -(np.roll(sh, -n, axis=3) * g42).sum() * st['temp'] / Nkall
    return out


def test_chi0_soc_matches_direct_sum(soc):
    st = soc
    chi = O.get_chi0_conv_soc(st['Gk'], *_args(st)).T
    ref = _chi0_direct(st)
    # the FFT result is thresholded at 1e-9 per component
    assert np.abs(chi - ref).max() < 2e-9
    assert np.abs(ref).max() > 0.1


def test_soc_vertex_identity(soc):
    st = soc
    chi = O.get_chi0_conv_soc(st['Gk'], *_args(st))
    V = O.get_V_soc_flex(chi, st['Vmat']).T
    A = np.asarray(chi).T
    Vf = np.asarray(st['Vmat']).T
    alt = -Vf @ np.linalg.inv(np.eye(len(Vf)) + A @ Vf)
    assert np.abs(V - alt).max() < 1e-12 * np.abs(V).max()


def test_sigma_soc_matches_direct_sum(soc):
    st = soc
    chi = O.get_chi0_conv_soc(st['Gk'], *_args(st))
    V = O.get_V_soc_flex(chi, st['Vmat'])
    a = (st['Gk'], V, st['Vmat'], st['kmap'], st['invk'], st['invs'], st['olist'], st['slist'], st['temp'],
         st['Nx'], st['Ny'], st['Nz'])
    s1 = O.calc_sigma_soc(*a)
    s2 = O.calc_sigma_soc_direct(*a)
    assert np.abs(s1 - s2).max() < 2e-9 and np.abs(s2).max() > 1e-2


def _soc_call(st, mu, fill, **kw):
    return O.mkself_soc(st['Vmat'], st['kmap'], st['invk'], st['invs'], st['olist'], st['slist'], st['hamk'],
                        st['eig'], st['uni'], mu, fill, st['temp'], st['Nw'], st['Nx'], st['Ny'], st['Nz'],
                        False, False, **kw)


def test_mkself_soc_fixed_point(soc):
    st = soc
    fill = 1.6
    sig, mu, info = _soc_call(st, 0.0, fill, eps=1e-9, pp=0.5, m_diis=1, scf_loop=200, return_info=True)
    assert info['esterr'][-1] < 1e-9
    # the loop map at the converged point, built from the pieces with the
    # mu of the last Green function
    G = O.gen_green(sig, st['hamk'], info['mu_file'], st['temp'])
    chi = O.get_chi0_conv_soc(G, *_args(st))
    V = O.get_V_soc_flex(chi, st['Vmat'])
    out = O.calc_sigma_soc(G, V, st['Vmat'], st['kmap'], st['invk'], st['invs'], st['olist'], st['slist'],
                           st['temp'], st['Nx'], st['Ny'], st['Nz']).T
    base = out[:, 0]
    out = out - 0.5 * (base + np.conj(np.swapaxes(base, 1, 2)))[:, None]
    assert np.abs(out - np.asarray(sig).T).max() < 1e-7 * np.abs(sig).max()
    assert np.abs(sig).max() > 1e-3
