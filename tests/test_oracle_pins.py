"""Pins of the oracle functions the reference's own tests leave open, by the
mathematics: an exact band-sum bubble for the multi-orbital tail-corrected
chi0, closed forms for the pairing vertex and for the Eliashberg operator
with a constant vertex, direct circular sums for the gap convolution, the
element-wise definition of the anomalous Green function, the full-zone
normalisation of the initial gap, and fixed-point residuals of the FLEX loop
(multi-orbital, and through the Stoner-rescale branch).  CPU only."""
import numpy as np
import pytest

import oracle.pyrpa_oracle as O
from tests import models as MD


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


# ----------------------------------------------------------------------------- chi0 tail
@pytest.mark.parametrize("norb", [2, 3])
def test_multiorbital_tail_chi0_converges_to_the_band_sum_at_second_order(norb):
    """The tail-corrected chi0 (conv[G] - conv[G0] + ref[G0], fchi.f90:852-892)
    of a hybridised multi-orbital G0 against the exact Matsubara bubble from
    a band sum: its error falls by more than 9x when Nw grows 4x (second
    order), the sharp-cutoff one by about 4x (first order), and the tail is
    the more accurate one (the reference's own criterion for one orbital,
    tests/test_rpa_flex.py:308-333, here for every orbital component)."""
    ml = [0, 1, 3]
    errs = {}
    for Nw in (16, 64):
        if norb == 2:
            st = MD.two_orbital(Nx=6, Ny=6, Nw=Nw, temp=0.1, mu=0.1, V=0.3)
        else:
            st = MD.random_model(3, 4, 4, 1, Nw, 0.1, 0.05, 0.0, 0.0, seed=5, real=True)
        ex = MD.bubble_exact(st, ml)
        args = (st['Smat'], st['Cmat'], st['Gk'], st['eig'], st['uni'], st['olist'], st['kmap'], st['invk'],
                st['mu'], st['temp'], st['Nx'], st['Ny'], st['Nz'])
        tail, _ = O.get_chi0_tail(*args)
        conv = O.get_chi0_conv(st['Gk'], st['kmap'], st['invk'], st['olist'], st['temp'], st['Nx'], st['Ny'],
                               st['Nz'])
        errs[Nw] = (np.abs(conv[:, :, ml] - ex).max(), np.abs(tail[:, :, ml] - ex).max())
    assert errs[16][1] / errs[64][1] > 9.0, errs
    assert 2.5 < errs[16][0] / errs[64][0] < 7.0, errs
    assert errs[64][1] < 0.2 * errs[64][0]


def test_band_sum_bubble_is_the_lindhard_function_for_one_orbital():
    st = MD.one_orbital(Nx=6, Ny=6, Nw=8, temp=0.1, mu=0.2)
    a = MD.bubble_exact(st, [0, 2])
    b = MD.lindhard_exact(st, [0, 2])
    assert np.abs(a[0, 0] - b.T).max() < 1e-14


# ----------------------------------------------------------------------------- V_Delta
def test_pairing_vertex_one_orbital_closed_form():
    """fUJ.f90:253-316 with S = C = U: singlet U + 1.5 U^2 chi/(1 - U chi)
    - 0.5 U^2 chi/(1 + U chi); triplet -0.5 U^2 chi/(1 - U chi) - 0.5 U^2
    chi/(1 + U chi)."""
    rng = np.random.default_rng(2)
    chi = 0.2 * (rng.standard_normal((1, 1, 5, 7)) + 1j * rng.standard_normal((1, 1, 5, 7)))
    U = 1.3
    S = C = np.array([[U]])
    x = chi[0, 0]
    vs = U + 1.5 * U * U * x / (1 - U * x) - 0.5 * U * U * x / (1 + U * x)
    vt = -0.5 * U * U * x / (1 - U * x) - 0.5 * U * U * x / (1 + U * x)
    assert np.abs(O.get_V_delta(chi, S, C, singlet=True)[0, 0] - vs).max() < 1e-14
    assert np.abs(O.get_V_delta(chi, S, C, singlet=False)[0, 0] - vt).max() < 1e-14
    # the FLEX vertex of the same chi (fUJ.f90:318-365 comment)
    vf = U + 1.5 * U * U * x / (1 - U * x) + 0.5 * U * U * x / (1 + U * x) - U * U * x
    assert np.abs(O.get_Vsigma_nosoc_flex(chi, S, C)[0, 0] - vf).max() < 1e-14


# ----------------------------------------------------------------------------- mkfk, mkdelta
def test_anomalous_green_function_elementwise():
    """F = -G Delta G^dagger per (k, i w), then the lower triangle copied
    conjugated to the upper one and the diagonal made real
    (feliash.f90:445-484), element by element."""
    rng = np.random.default_rng(7)
    No, Nw, Nk = 3, 2, 3
    G = rng.standard_normal((No, No, Nw, Nk)) + 1j * rng.standard_normal((No, No, Nw, Nk))
    D = rng.standard_normal((No, No, Nw, Nk)) + 1j * rng.standard_normal((No, No, Nw, Nk))
    F = O.mkfk_trs_nsoc(G, D)
    for k in range(Nk):
        for w in range(Nw):
            g = G[:, :, w, k].T       # g[a, b] = G_ab (Python Gk[b, a, w, k])
            d = D[:, :, w, k].T
            f = np.zeros((No, No), dtype=complex)
            for a in range(No):
                for b in range(No):
                    for c in range(No):
                        for e in range(No):
                            f[a, b] -= g[a, c] * d[c, e] * np.conj(g[b, e])
            for a in range(No):
                for b in range(No):
                    if a == b:
                        want = f[a, a].real
                    elif a < b:
                        want = np.conj(f[b, a])
                    else:
                        want = f[a, b]
                    assert abs(F[b, a, w, k] - want) < 1e-12


def _mkdelta_direct(fk, Vd, S, C, kmap, invk, prt, olist, Nx, Ny, Nz, singlet):
    """The gap convolution of feliash.f90:486-635 as an explicit circular
    sum over the full (k, tau-slot) grid instead of FFTs."""
    Ff = np.asarray(fk).T
    Vf = np.asarray(Vd).T
    Sf, Cf = np.asarray(S).T, np.asarray(C).T
    ol = O.olist_f(olist)
    Nk, Nw, No, _ = Ff.shape
    M = 2 * Nw
    sgn = 1.0 if singlet else -1.0
    nd = np.zeros((Nk, Nw, No, No), dtype=complex)
    d = np.nonzero(invk[:, 1] == 0)[0]
    for l in range(len(ol)):
        for m in range(len(ol)):
            if ol[l, 0] < ol[m, 1]:
                continue
            v = O._full_boson(Vf, m, l, 0.5 * (Sf[m, l] + sgn * Cf[m, l]), kmap, invk, Nx, Ny, Nz)
            f = O._full_anom(Ff, ol[l, 1], ol[m, 0], prt, sgn, sgn, kmap, invk, Nx, Ny, Nz)
            for i in d:
                kx, ky, kz = kmap[i]
                for n in range(Nw):
                    acc = 0j
                    for qx in range(Nx):
                        for qy in range(Ny):
                            for qz in range(Nz):
                                for s in range(M):
                                    acc += v[qx, qy, qz, s] * f[(kx - qx) % Nx, (ky - qy) % Ny, (kz - qz) % Nz,
                                                               (n - s) % M]
                    nd[invk[i, 0] - 1, n, ol[l, 0], ol[m, 1]] += acc
    for l in range(No):
        for m in range(l, No):
            if l != m:
                nd[:, :, l, m] = np.conj(nd[:, :, m, l])
            else:
                nd[:, :, l, l] = nd[:, :, l, l].real
    return nd


@pytest.mark.parametrize("singlet", [True, False])
def test_gap_convolution_matches_direct_sum(singlet):
    st = MD.two_orbital(Nx=2, Ny=2, Nw=3, temp=0.2, mu=0.1, U=0.8, J=0.1)
    chi = O.get_chi0_conv(st['Gk'], st['kmap'], st['invk'], st['olist'], st['temp'], 2, 2, 1)
    Vd = O.get_V_delta(chi, st['Smat'], st['Cmat'], singlet=singlet)
    rng = np.random.default_rng(9)
    D = rng.standard_normal(st['Gk'].shape) + 1j * rng.standard_normal(st['Gk'].shape)
    fk = O.mkfk_trs_nsoc(st['Gk'], D)
    prt = np.array([1.0, -1.0])
    a = O.mkdelta_nsoc(fk, Vd, st['Smat'], st['Cmat'], st['kmap'], st['invk'], prt, st['olist'], 2, 2, 1,
                       singlet=singlet)
    b = _mkdelta_direct(fk, Vd, st['Smat'], st['Cmat'], st['kmap'], st['invk'], prt, st['olist'], 2, 2, 1, singlet)
    assert np.abs(a - b).max() < 1e-12 * max(1.0, np.abs(b).max())


@pytest.mark.parametrize("arnoldi_m", [0, 10])
def test_eliashberg_with_a_constant_vertex_has_the_closed_form_eigenvalue(arnoldi_m):
    """chi0 = 0 leaves the constant singlet vertex V_Delta = (S + C)/2 = U,
    the nu -> inf slot included.  The operator then maps every gap to the
    constant U (T/Nkall) sum over the full zone and all 2Nw Matsubara slots
    of F = -|G|^2 Delta, so its one nonzero eigenvalue is
    lambda = -2 U (T/Nkall) sum_{k in BZ} sum_{n < Nw} |G(k, i w_n)|^2
    (factor 2: the negative frequencies, F(k,-w) = F*(k,w)).  Attractive U."""
    st = MD.one_orbital(Nx=4, Ny=4, Nw=8, temp=0.05, mu=-0.2, tx=1.0, ty=1.0)
    U = -1.0
    S = C = np.array([[U]])
    Nk = len(st['klist'])
    chi = np.zeros((1, 1, 8, Nk), dtype=complex)
    w = O.kweight(st['invk'], Nk)
    lam_exact = -2.0 * U * st['temp'] / 16 * (w[None, :] * np.abs(st['Gk'][0, 0]) ** 2).sum()
    rng = np.random.default_rng(1)
    init = rng.uniform(-1, 1, (1, Nk))
    gap, lam = O.linearized_eliashberg(chi, st['Gk'], st['uni'], init, S, C, st['olist'], st['plist'], st['kmap'],
                                       st['invk'], 4, 4, 1, st['temp'], 2, eps=1e-12, itemax=50,
                                       arnoldi_m=arnoldi_m)
    assert abs(lam - lam_exact) < 1e-12 * lam_exact
    g = gap[0, 0]
    assert np.abs(g - g.flat[0]).max() < 1e-10 * np.abs(g).max()   # constant eigenvector


def test_initial_gap_normalised_over_the_full_zone():
    """get_initial_delta (feliash.f90:637-731): Delta0(k) = U diag(d) U^dagger
    (s: the identity), the same at every w, and 2 sum over the full zone, w
    and orbitals of |Delta0|^2 = 1, each full-zone point counted through its
    irreducible partner."""
    st = MD.random_model(3, 4, 4, 1, 3, 0.1, 0.0, 0.0, 0.0, seed=4)
    Nk = len(st['klist'])
    rng = np.random.default_rng(6)
    init = rng.uniform(-1, 1, (3, Nk))
    for gs in (0, 1, -1):
        D = O.get_initial_delta(init, st['uni'], st['kmap'], st['invk'], 3, gs)
        tot = 0.0
        for i in range(len(st['kmap'])):
            tot += (np.abs(D[:, :, :, st['invk'][i, 0] - 1]) ** 2).sum()
        assert abs(2.0 * tot - 1.0) < 1e-12
        scale = None
        for k in range(Nk):
            if gs == 0:
                want = np.eye(3)
            else:
                U = st['uni'][k].T      # U[orb, band]
                want = U @ np.diag(init[:, k]) @ np.conj(U.T)
            got = D[:, :, 0, k].T
            if scale is None:
                scale = np.abs(got).max() / np.abs(want).max()
            assert np.abs(got - scale * want).max() < 1e-12
            assert np.abs(D[:, :, 2, k] - D[:, :, 0, k]).max() == 0.0


# ----------------------------------------------------------------------------- FLEX loop
def _flex_map(st, sig, mu, S, C, sub_sigma=1, rescale=False):
    """One FLEX map from (Sigma, mu) with the direct-sum chi0 and Sigma:
    G(Sigma, mu) -> chi0 (O(Nk^2) sum, thresholded) -> [rescale] -> V_sigma ->
    Sigma (direct circular sum, thresholded) -> static subtraction."""
    G = O.gen_green(sig, st['hamk'], mu, st['temp']) if sig is not None else st['Gk']
    nx, ny, nz = st['Nx'], st['Ny'], st['Nz']
    chi = O.get_chi0_sum(G, st['klist'], st['kmap'], st['invk'], st['olist'], st['temp'], nx, ny, nz)
    chi = O._threshold(np.asarray(chi).T).T
    stoner = O.ckchi(chi, S, C)
    if rescale and stoner >= 1.0:
        chi = chi * ((1.0 - 1.0e-4) / stoner)
    V = O.get_Vsigma_nosoc_flex(chi, S, C)
    s2 = np.asarray(O.calc_sigma_direct(G, V, S, C, st['kmap'], st['invk'], st['olist'], st['temp'], nx, ny,
                                        nz)).T.copy()
    s2 = O._threshold(s2)
    if sub_sigma == 1:
        base = s2[:, 0]
        sub = 0.5 * (base + np.conj(np.swapaxes(base, 1, 2)))
        s2 = s2 - sub[:, None]
    return np.ascontiguousarray(s2.T), chi, stoner


def test_two_orbital_mkself_is_a_fixed_point_of_direct_components():
    """The converged 2-orbital Sigma (Kanamori U, J; sub_sigma = 1) reproduces
    itself through one FLEX map built from direct sums, and its mu gives the
    target filling through the resolvent trace."""
    st = MD.two_orbital(Nx=4, Ny=4, Nw=8, temp=0.1, mu=0.0)
    S, C = O.get_scmat(st['olist'], st['site'], 0.8, 0.1)
    sig, mu, info = O.mkself(S, C, st['kmap'], st['invk'], st['olist'], st['hamk'], st['eig'], st['uni'], 0.0,
                             0.9, st['temp'], 8, 4, 4, 1, False, False, eps=1e-11, pp=0.5, m_diis=1,
                             scf_loop=300, return_info=True)
    assert info['esterr'][-1] < 1e-11
    s2, _, _ = _flex_map(st, sig, mu, S, C)
    assert rel(s2, sig) < 1e-8
    G = O.gen_green(sig, st['hamk'], mu, st['temp'])
    G0 = O.gen_Green0(st['eig'], st['uni'], mu, st['temp'], 8)
    Gf, G0f = np.asarray(G).T, np.asarray(G0).T
    n = (O.get_ffermi(st['eig'], mu, st['temp']).sum()
         + 2 * st['temp'] * np.trace(Gf - G0f, axis1=2, axis2=3).real.sum()) / len(st['klist'])
    assert abs(n - 0.9) < 1e-8


def test_stoner_rescale_branch():
    """sw_rescale with a Stoner factor >= 1 (fself.f90:371-380): chi0 is
    scaled to (1 - 1e-4)/S before the vertex, so the rescaled chi0's Stoner
    factor is 1 - 1e-4; one cold-start iteration with linear mixing returns
    pp times the directly computed map."""
    st = MD.one_orbital(Nx=4, Ny=4, Nw=16, temp=0.1, mu=0.0, tx=1.0, ty=1.0)
    S = C = np.array([[5.0]])
    _, chi, stoner = _flex_map(st, None, 0.0, S, C)
    assert stoner > 1.0
    assert abs(O.ckchi(chi * ((1.0 - 1.0e-4) / stoner), S, C) - (1.0 - 1.0e-4)) < 1e-12
    sig, mu, info = O.mkself(S, C, st['kmap'], st['invk'], st['olist'], st['hamk'], st['eig'], st['uni'], 0.0,
                             0.5, st['temp'], 16, 4, 4, 1, False, False, eps=0.0, pp=0.5, m_diis=1, scf_loop=1,
                             sw_rescale=True, return_info=True)
    assert info['stoner'][0] > 1.0
    s_map, _, _ = _flex_map(st, None, 0.0, S, C, rescale=True)
    assert rel(sig, 0.5 * s_map) < 1e-12
    # without the rescale the same iteration is a different map
    s_plain, _, _ = _flex_map(st, None, 0.0, S, C, rescale=False)
    assert rel(0.5 * s_plain, sig) > 1e-3
