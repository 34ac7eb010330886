"""Pins the CPU oracle before it is trusted as the parity reference.

Each check is independent of the restatement it pins: closed forms, golden
values held by the reference's own tests (tests/test_rpa_flex.py,
tests/test_tools.py), direct O(N^2) sums in place of the FFT convolutions,
and invariants.  All CPU, seconds in total."""
import numpy as np
import pytest

import oracle.pyrpa_oracle as O
from tests import models as MD


# ----------------------------------------------------------------------------- k-mesh
@pytest.mark.parametrize("mesh", [(2, 2, 1), (4, 4, 1), (8, 8, 1), (4, 4, 4), (3, 3, 1),
                                  (4, 3, 1), (4, 4, 3), (6, 4, 2), (5, 5, 3)])
def test_irreducible_mesh_pairs_every_point(mesh):
    Nx, Ny, Nz = mesh
    klist, kmap, invk = O.gen_irr_k_TRS(Nx, Ny, Nz)
    N = np.array(mesh)
    kc = np.rint(klist * N).astype(int)
    assert len(klist) == O._irr_nk(Nx, Ny, Nz)
    # no irreducible point twice, and never both k and -k (except TRIM)
    keys = {tuple(k) for k in kc}
    assert len(keys) == len(kc)
    for k in kc:
        mk = tuple((-k) % N)
        assert mk == tuple(k) or mk not in keys
    for i in range(len(kmap)):
        irr = invk[i, 0] - 1
        k = kmap[i]
        if invk[i, 1] == 0:
            assert tuple(kc[irr]) == tuple(k)
        else:
            assert tuple((-kc[irr]) % N) == tuple(k)
        mk = kmap[invk[i, 2] - 1]
        assert tuple(mk) == tuple((-k) % N)
    assert O.kweight(invk, len(klist)).sum() == Nx * Ny * Nz


def test_irreducible_mesh_order_4x4():
    klist, _, _ = O.gen_irr_k_TRS(4, 4, 1)
    want = [(0, 0), (1, 0), (2, 0), (0, 1), (1, 1), (2, 1), (0, 2), (1, 2), (2, 2), (1, 3)]
    assert np.allclose(klist[:, :2] * 4, want)


# ----------------------------------------------------------------------------- vertices
def test_uniform_sc_matrix_two_orbital_reference_values():
    olist, site = O.get_chi_orb_list(2, [2])
    S, C = O.get_scmat(olist, site, 1.0, 0.2)
    es = [[1.0, 0, 0, 0.2], [0, 0.6, 0.2, 0], [0, 0.2, 0.6, 0], [0.2, 0, 0, 1.0]]
    ec = [[1.0, 0, 0, 1.0], [0, -0.2, 0.2, 0], [0, 0.2, -0.2, 0], [1.0, 0, 0, 1.0]]
    assert np.allclose(S, es) and np.allclose(C, ec)


def test_orbital_dependent_vertices_reference_values():
    olist, site = O.get_chi_orb_list(2, [2])
    S, C = O.get_scmat_orb(olist, site, np.full((2, 2), 1.0), np.full((2, 2), 0.2))
    es = [[1.0, 0, 0, 0.2], [0, 1.0, 0.2, 0], [0, 0.2, 1.0, 0], [0.2, 0, 0, 1.0]]
    ec = [[1.0, 0, 0, 1.8], [0, -0.6, 0.2, 0], [0, 0.2, -0.6, 0], [1.8, 0, 0, 1.0]]
    assert np.allclose(S, es) and np.allclose(C, ec)


def test_chi_map_pairs_are_exchange_partners():
    olist, _ = O.get_chi_orb_list(3, [3])
    ol = O.olist_f(olist)
    cmap, irr = O.get_chi_map(olist)
    assert len(irr) == 9 * 10 // 2
    for m1 in range(9):
        for l1 in range(9):
            m2, l2 = cmap[m1, l1]
            # chi_{1234} = chi*_{4321}: (l1=(a,b), m1=(c,d)) <-> (l2=(d,c), m2=(b,a))
            assert tuple(ol[l2]) == (ol[m1, 1], ol[m1, 0])
            assert tuple(ol[m2]) == (ol[l1, 1], ol[l1, 0])


# ----------------------------------------------------------------------------- RPA algebra
def test_scalar_rpa_chis_matches_closed_form():
    chi0 = np.array([[[0.20 + 0.10j]], [[0.05 + 0.02j]]])
    U = 1.2
    chis = O.get_chis(chi0, np.array([[U]]))
    assert np.allclose(chis[:, 0, 0], chi0[:, 0, 0] / (1 - U * chi0[:, 0, 0]))


def test_chis_chic_reduce_to_bare_chi_when_vertices_are_zero():
    chi = np.zeros((2, 2, 1, 3), dtype=complex)
    chi[:, :, 0, 0] = [[0.10, 0.02j], [-0.02j, 0.20]]
    chi[:, :, 0, 1] = [[0.05, 0.01], [0.03, 0.04]]
    chis, chic = O.get_chis_chic(chi, np.zeros((2, 2)), np.zeros((2, 2)))
    assert np.allclose(chis, chi[:, :, 0, :]) and np.allclose(chic, chi[:, :, 0, :])


def test_vsigma_one_orbital_closed_form_and_zero_vertex():
    rng = np.random.default_rng(1)
    chi = (rng.uniform(0, 0.3, (1, 1, 5, 7)) + 1j * rng.uniform(-0.05, 0.05, (1, 1, 5, 7)))
    U = 1.3
    V = O.get_Vsigma_nosoc_flex(chi, np.array([[U]]), np.array([[U]]))
    x = chi[0, 0]
    want = U + 1.5 * U * U * x / (1 - U * x) + 0.5 * U * U * x / (1 + U * x) - U * U * x
    assert np.allclose(V[0, 0], want, rtol=1e-13, atol=1e-14)
    assert np.allclose(O.get_Vsigma_nosoc_flex(chi, np.zeros((1, 1)), np.zeros((1, 1))), 0)


def test_vsigma_multiorbital_resolvent_identity():
    """V = 1.5 S (I-AS)^-1 - 0.5 C (I+AC)^-1 - 0.25 (S+C) A (S+C), an algebraic
    rearrangement of the reference's X_s/X_c form (fUJ.f90:356)."""
    olist, site = O.get_chi_orb_list(2, [2])
    S, C = O.get_scmat(olist, site, 1.0, 0.2)
    rng = np.random.default_rng(2)
    chi = 0.1 * (rng.standard_normal((4, 4, 3, 5)) + 1j * rng.standard_normal((4, 4, 3, 5)))
    V = O.get_Vsigma_nosoc_flex(chi, S, C)
    A = np.asarray(chi).T
    Sf, Cf = S.T, C.T
    I = np.eye(4)
    want = (1.5 * Sf @ np.linalg.inv(I - A @ Sf) - 0.5 * Cf @ np.linalg.inv(I + A @ Cf)
            - 0.25 * (Sf + Cf) @ A @ (Sf + Cf))
    assert np.abs(np.asarray(V).T - want).max() < 1e-12


# ----------------------------------------------------------------------------- Green functions
def test_green0_noninteracting_inverse_for_one_orbital():
    st = MD.one_orbital(Nx=2, Ny=2, Nw=5, temp=0.07)
    wn = (2 * np.arange(5) + 1) * np.pi * 0.07
    assert np.allclose(st['Gk'][0, 0], 1.0 / (1j * wn[:, None] - st['eig'][:, 0][None, :]))


def test_two_orbital_bands_closed_form():
    st = MD.two_orbital(Nx=4, Ny=4)
    k = st['klist']
    cx, cy = np.cos(2 * np.pi * k[:, 0]), np.cos(2 * np.pi * k[:, 1])
    e1, e2 = -2 * (cx + cy), -2 * 0.5 * (cx + cy) + 0.4
    r = np.sqrt(((e1 - e2) / 2) ** 2 + 0.2 ** 2)
    want = np.sort(np.stack([(e1 + e2) / 2 - r, (e1 + e2) / 2 + r], 1), axis=1)
    assert np.allclose(st['eig'], want)


def test_dyson_inverse_identity_and_causality():
    st = MD.random_model(3, 4, 4, 1, 8, 0.1, 0.0, 0.0, 0.0, seed=3)
    rng = np.random.default_rng(4)
    sig = 0.05 * (rng.standard_normal((3, 3, 8, len(st['klist']))) - 1j * np.abs(rng.standard_normal((3, 3, 8, len(st['klist'])))))
    Gi = O.gen_green_inv(sig, st['hamk'], 0.1, 0.1)
    G = O.getinv(Gi)
    prod = np.asarray(G).T @ np.asarray(Gi).T
    assert np.abs(prod - np.eye(3)).max() < 1e-12
    G0 = MD.G0_at(st, 8, mu=0.1)
    assert (np.einsum('aawk->wk', G0).imag < 0).all()


# ----------------------------------------------------------------------------- chi0
@pytest.mark.parametrize("builder", ["one", "two", "three_3d"])
def test_chi0_fft_matches_direct_sum(builder):
    if builder == "one":
        st = MD.one_orbital(Nx=4, Ny=4, Nw=8, temp=0.1, mu=0.2)
    elif builder == "two":
        st = MD.two_orbital(Nx=4, Ny=2, Nw=4, temp=0.15, mu=-0.1)
    else:
        st = MD.random_model(3, 2, 2, 2, 3, 0.2, 0.05, 0, 0, seed=5)
    a = O.get_chi0_conv(st['Gk'], st['kmap'], st['invk'], st['olist'], st['temp'], st['Nx'], st['Ny'], st['Nz'])
    b = O.get_chi0_sum(st['Gk'], st['klist'], st['kmap'], st['invk'], st['olist'], st['temp'],
                       st['Nx'], st['Ny'], st['Nz'])
    b = O._threshold(np.asarray(b).T).T
    assert np.abs(a - b).max() < 1e-13 * max(1, np.abs(b).max())


def test_chi0_tail_matches_lindhard_and_converges_second_order():
    """tests/test_rpa_flex.py:308-333 against the oracle."""
    st = MD.one_orbital(Nx=8, Ny=8, Nw=8, temp=0.04, mu=0.3)
    ml = [0, 1, 2, 3]
    ex = MD.lindhard_exact(st, ml)
    errs = {}
    for Nw in (32, 128):
        G = MD.G0_at(st, Nw)
        co = O.get_chi0_conv(G, st['kmap'], st['invk'], st['olist'], st['temp'], 8, 8, 1)
        cn, _ = O.get_chi0_tail(st['Smat'], st['Cmat'], G, st['eig'], st['uni'], st['olist'],
                                st['kmap'], st['invk'], st['mu'], st['temp'], 8, 8, 1)
        eo = max(np.abs(co[0, 0, m] - ex[:, j]).max() for j, m in enumerate(ml))
        en = max(np.abs(cn[0, 0, m] - ex[:, j]).max() for j, m in enumerate(ml))
        errs[Nw] = (eo, en)
    assert 2.5 < errs[32][0] / errs[128][0] < 7.0
    assert errs[32][1] / errs[128][1] > 9.0
    assert errs[128][1] < 0.35 * errs[128][0]


def test_chi0_tail_exact_at_q0():
    """tests/test_rpa_flex.py:336-354: tail chi0(q=0,nu=0) = sum f(1-f)/T/N to 1e-10."""
    st = MD.one_orbital(Nx=8, Ny=8, Nw=32, temp=0.04, mu=0.3)
    cn, _ = O.get_chi0_tail(st['Smat'], st['Cmat'], st['Gk'], st['eig'], st['uni'], st['olist'],
                            st['kmap'], st['invk'], st['mu'], st['temp'], 8, 8, 1)
    ex0 = MD.lindhard_exact(st, [0])[0, 0]
    assert abs(cn[0, 0, 0, 0] - ex0) < 1e-10


def test_stoner_factor_zero_without_interaction_and_psd_statics():
    """tests/test_rpa_flex.py:157-171 and tests/test_tools.py:184-199."""
    st = MD.one_orbital(Nx=2, Ny=2, Nw=4)
    chi, stoner = O.get_chi0(st['Smat'], st['Cmat'], st['Gk'], st['olist'], st['kmap'], st['invk'],
                             st['temp'], 2, 2, 1)
    assert abs(stoner) < 1e-12
    st = MD.two_orbital(Nx=4, Ny=4, Nw=64, temp=0.1, U=0.5, J=0.1)
    chi, stoner = O.get_chi0(st['Smat'], st['Cmat'], st['Gk'], st['olist'], st['kmap'], st['invk'],
                             st['temp'], 4, 4, 1)
    assert 0.0 < stoner < 1.0
    chis, chic = O.get_chis_chic(chi, st['Smat'], st['Cmat'])
    for c in (chis, chic):
        M = np.transpose(c, (2, 0, 1))
        assert np.abs(M - np.conj(np.swapaxes(M, 1, 2))).max() < 1e-10
        assert np.linalg.eigvalsh(0.5 * (M + np.conj(np.swapaxes(M, 1, 2)))).min() > -1e-10


# ----------------------------------------------------------------------------- Sigma
def test_calc_sigma_fft_matches_direct_sum():
    st = MD.two_orbital(Nx=2, Ny=2, Nw=3, temp=0.2, mu=0.1, U=0.8, J=0.1)
    chi = O.get_chi0_conv(st['Gk'], st['kmap'], st['invk'], st['olist'], st['temp'], 2, 2, 1)
    V = O.get_Vsigma_nosoc_flex(chi, st['Smat'], st['Cmat'])
    a = O.calc_sigma(st['Gk'], V, st['Smat'], st['Cmat'], st['kmap'], st['invk'], st['olist'], st['temp'], 2, 2, 1)
    b = O.calc_sigma_direct(st['Gk'], V, st['Smat'], st['Cmat'], st['kmap'], st['invk'], st['olist'], st['temp'], 2, 2, 1)
    b = O._threshold(np.asarray(b).T).T
    assert np.abs(a - b).max() < 1e-13


# ----------------------------------------------------------------------------- FLEX loop
def _flex_one(U=1.0, eps=1e-10):
    st = MD.one_orbital(Nx=4, Ny=4, Nw=16, temp=0.1, U=U)
    sig, mu, info = O.mkself(st['Smat'], st['Cmat'], st['kmap'], st['invk'], st['olist'], st['hamk'],
                             st['eig'], st['uni'], 0.0, 0.4, st['temp'], st['Nw'], 4, 4, 1, False, False,
                             eps=eps, pp=0.5, m_diis=1, scf_loop=200, return_info=True)
    return st, sig, mu, info


def test_renew_mu_is_unreachable_when_sigma_vanishes():
    """Reference quirk: with Sigma == 0 every |Sigma| < 1e-10, so esterr = 0 and
    renew_mu's tolerance 0.1*esterr is zero (fself.f90:552-556); its Brent loop
    can only end at its 100-step STOP.  The oracle raises instead."""
    with pytest.raises(RuntimeError):
        _flex_one(U=0.0)


def test_mkself_converged_sigma_is_a_fixed_point_of_direct_components():
    """Independent check of the whole loop: recompute one FLEX map from the
    returned (Sigma, mu) with the direct-sum chi0 and Sigma and the resolvent
    n(mu); the result must reproduce Sigma and the target filling."""
    st, sig, mu, info = _flex_one(U=1.5, eps=1e-11)
    assert info['esterr'][-1] < 1e-11
    G = O.gen_green(sig, st['hamk'], mu, st['temp'])
    chi = O.get_chi0_sum(G, st['klist'], st['kmap'], st['invk'], st['olist'], st['temp'], 4, 4, 1)
    chi = O._threshold(np.asarray(chi).T).T
    V = O.get_Vsigma_nosoc_flex(chi, st['Smat'], st['Cmat'])
    s2 = np.asarray(O.calc_sigma_direct(G, V, st['Smat'], st['Cmat'], st['kmap'], st['invk'],
                                        st['olist'], st['temp'], 4, 4, 1)).T.copy()
    s2 = O._threshold(s2)
    s2[:, :, 0, 0] -= s2[:, 0, 0, 0].real[:, None]
    assert np.abs(s2 - np.asarray(sig).T).max() < 1e-8 * np.abs(sig).max()
    # filling from the resolvent trace, not from eigenvalues
    Gf = np.asarray(G).T
    G0f = np.asarray(O.gen_Green0(st['eig'], st['uni'], mu, st['temp'], 16)).T
    n = (O.get_ffermi(st['eig'], mu, st['temp']).sum()
         + 2 * st['temp'] * np.trace(Gf - G0f, axis1=2, axis2=3).real.sum()) / len(st['klist'])
    assert abs(n - 0.4) < 1e-8


# ----------------------------------------------------------------------------- Eliashberg
def test_linearized_eliashberg_zero_interaction_has_zero_lambda():
    """tests/test_rpa_flex.py:187-205 against the oracle (Arnoldi m=2)."""
    st = MD.one_orbital(Nx=2, Ny=2, Nw=4)
    chi, _ = O.get_chi0(st['Smat'], st['Cmat'], st['Gk'], st['olist'], st['kmap'], st['invk'],
                        st['temp'], 2, 2, 1)
    init = O.gap_symms(st['klist'], 1, 0)
    gap, lam = O.linearized_eliashberg(chi, st['Gk'], st['uni'], init, st['Smat'], st['Cmat'],
                                       st['olist'], st['plist'], st['kmap'], st['invk'], 2, 2, 1,
                                       st['temp'], 0, itemax=3, arnoldi_m=2)
    assert abs(lam) < 1e-12 and np.isfinite(gap).all()


def test_eliashberg_power_and_arnoldi_agree_with_dense_operator():
    """lambda from power iteration and from Arnoldi (m=Nk*Nw: exact Krylov)
    is an eigenvalue of the dense real matrix of the same operator, and the
    d-wave gap is odd under kx<->ky."""
    st = MD.one_orbital(Nx=4, Ny=4, Nw=4, temp=0.05, mu=-0.2, tx=1.0, ty=1.0, U=2.0)
    chi, _ = O.get_chi0(st['Smat'], st['Cmat'], st['Gk'], st['olist'], st['kmap'], st['invk'],
                        st['temp'], 4, 4, 1)
    init = O.gap_symms(st['klist'], 1, 1)
    args = (chi, st['Gk'], st['uni'], init, st['Smat'], st['Cmat'], st['olist'], st['plist'],
            st['kmap'], st['invk'], 4, 4, 1, st['temp'], 1)
    gp, lp = O.linearized_eliashberg(*args, eps=1e-12, itemax=2000, arnoldi_m=0)
    ga, la = O.linearized_eliashberg(*args, arnoldi_m=40)
    # dense operator on real diagonal gaps
    Nk, Nw = len(st['klist']), 4
    Vd = O.get_V_delta(chi, st['Smat'], st['Cmat'])
    K = np.zeros((Nk * Nw, Nk * Nw))
    for c in range(Nk * Nw):
        D = np.zeros((1, 1, Nw, Nk), dtype=complex)
        D[0, 0].T.flat[c] = 1.0
        fk = O.mkfk_trs_nsoc(st['Gk'], D)
        nd = O.mkdelta_nsoc(fk, Vd, st['Smat'], st['Cmat'], st['kmap'], st['invk'], st['plist'],
                            st['olist'], 4, 4, 1) * (st['temp'] / 16)
        K[:, c] = nd[:, :, 0, 0].real.ravel()
    ev = np.linalg.eigvals(K)
    assert np.min(np.abs(ev - lp)) < 1e-8
    assert np.min(np.abs(ev - la)) < 1e-8
    assert abs(la - lp) < 1e-6
    # d-wave: Delta(kx,ky) = -Delta(ky,kx) on the irreducible list
    kc = np.rint(st['klist'][:, :2] * 4).astype(int)
    g = ga[0, 0, 0]
    pos = {tuple(k): i for i, k in enumerate(kc)}
    for i, (x, y) in enumerate(kc):
        j = pos.get((y, x)) or pos.get(((-y) % 4, (-x) % 4))
        if j is not None:
            assert abs(g[i] + g[j]) < 1e-8 * np.abs(g).max()
