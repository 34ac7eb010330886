"""GPU parity of the spinful (spin-orbit) path against the oracle through the
C-ABI: get_chi0_soc, the vertex of fUJ.f90:367-399, get_chis_chic_soc and
mkself_soc (fchi0withsoc.f90, fselfwithsoc.f90).  The oracle is pinned in
tests/test_soc_oracle.py.  FP64 on both sides in a different summation order:
chi0 and the vertex to 1e-11 of their largest element, the loop to the
iteration count of the oracle, Sigma to 1e-8 relative and mu to 1e-8.  The
FP32-transform loop lands within 1e-5 of the FP64 one."""
import numpy as np
import pytest

import oracle.pyrpa_oracle as O
from tests import models as MD

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def F():
    import pyrpa_amd.flibs as F
    F.set_precision(0)
    return F


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


def _model(name):
    if name == "degenerate_2orb":
        base = MD.two_orbital(Nx=4, Ny=4, Nw=8, temp=0.1, mu=0.1, V=0.3)
        return MD.spinful(base, U=0.8, J=0.15)
    if name == "soc_2orb":
        base = MD.two_orbital(Nx=4, Ny=4, Nw=8, temp=0.1, mu=0.1, V=0.3)
        return MD.spinful(base, lam=0.4, alpha=0.15, U=0.8, J=0.15)
    if name == "soc_2orb_6x4":
        base = MD.two_orbital(Nx=6, Ny=4, Nw=16, temp=0.08, mu=0.0, V=0.2)
        return MD.spinful(base, lam=0.3, alpha=0.1, U=0.6, J=0.1)
    base = MD.one_orbital(Nx=6, Ny=6, Nw=16, temp=0.1, mu=0.2)
    return MD.spinful(base, alpha=0.2, U=1.2)


CASES = ["degenerate_2orb", "soc_2orb", "soc_2orb_6x4", "rashba_1orb"]


@pytest.mark.parametrize("name", CASES)
def test_chi0_soc_matches_oracle(F, name):
    st = _model(name)
    args = (st['Vmat'], st['Gk'], st['olist'], st['slist'], st['kmap'], st['invk'], st['invs'], st['temp'],
            st['Nx'], st['Ny'], st['Nz'])
    a = F.get_chi0_soc(*args)
    b = O.get_chi0_soc(*args)
    assert rel(a[0], b[0]) < 1e-11
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2]) and np.array_equal(a[3], b[3])
    # the vertex and the channel susceptibilities from the same chi0
    assert rel(F.get_V_soc_flex(b[0], st['Vmat']), O.get_V_soc_flex(b[0], st['Vmat'])) < 1e-11
    x = F.get_chis_chic_soc(b[0], st['Vmat'], st['olist'], st['slist'], st['invs'])
    y = O.get_chis_chic_soc(b[0], st['Vmat'], st['olist'], st['slist'], st['invs'])
    for u, v in zip(x, y):
        assert rel(u, v) < 1e-11


def test_vmat_soc_matches_oracle(F):
    for No in (2, 4, 6):
        slist, invs = O.soc_spin_lists(No)
        olist, site = O.get_chi_orb_list(No, [No])
        V = F.gen_Vmatrix(olist, slist, site, invs, 1.1, 0.25)
        assert np.array_equal(V, O.get_vmat_soc(olist, slist, site, invs, 1.1, 0.25))


def _soc_args(st, fill):
    return (st['Vmat'], st['kmap'], st['invk'], st['invs'], st['olist'], st['slist'], st['hamk'], st['eig'],
            st['uni'], 0.0, fill, st['temp'], st['Nw'], st['Nx'], st['Ny'], st['Nz'], False, False)


@pytest.mark.parametrize("name,fill,kw", [("soc_2orb", 1.6, {}), ("degenerate_2orb", 1.8, dict(sub_sigma=2)),
                                          ("rashba_1orb", 0.9, dict(m_diis=3))])
def test_mkself_soc_matches_oracle(F, name, fill, kw):
    st = _model(name)
    opts = dict(eps=1e-10, pp=0.5, m_diis=1, scf_loop=200)
    opts.update(kw)
    sa, ma = F.mkself_soc(*_soc_args(st, fill), **opts)
    sb, mb, info = O.mkself_soc(*_soc_args(st, fill), return_info=True, **opts)
    print(f"{name}: {info['iters']} iterations, max|Sigma| {np.abs(sb).max():.3e}, mu {mb:.8f}")
    assert F.flex_stats()['iterations'] == info['iters']
    assert rel(sa, sb) < 1e-8
    assert abs(ma - mb) < 1e-8


def test_mkself_soc_fp32_transforms(F):
    st = _model("soc_2orb_6x4")
    opts = dict(eps=1e-8, pp=0.5, m_diis=1, scf_loop=300)
    s64, m64 = F.mkself_soc(*_soc_args(st, 1.7), **opts)
    try:
        F.set_precision(2)
        s32, m32 = F.mkself_soc(*_soc_args(st, 1.7), **opts)
    finally:
        F.set_precision(0)
    print(f"FP32 transforms: max|dSigma|/max|Sigma| {rel(s32, s64):.2e}, |dmu| {abs(m32 - m64):.2e}")
    assert rel(s32, s64) < 1e-5 and abs(m32 - m64) < 1e-5


def _eliash_args(st, chi, sg, sg2, isc, gs):
    init = O.gap_symms(st['klist'], st['Norb'], gs)
    return (chi, st['Gk'], st['uni'], init, st['Vmat'], sg, sg2, np.ones(st['Norb']), st['slist'], st['olist'],
            st['kmap'], st['invk'], st['invs'], isc, st['Nx'], st['Ny'], st['Nz'], st['temp'], gs)


@pytest.mark.parametrize("name", ["soc_2orb", "rashba_1orb"])
@pytest.mark.parametrize("gs", [1, -1])
def test_lin_eliash_soc_operator_matches_oracle(F, name, gs):
    """The gap operator of lin_eliash_soc (feliashwithsoc.f90:426-663) step by
    step: the power method stopped after 1, 2 and 5 steps returns the
    normalised last iterate and its Rayleigh quotient; the quotient agrees
    with the oracle's to 1e-10, the iterate to 1e-7 (components the steps
    shrink keep the rounding of the first step)."""
    st = _model(name)
    chi, sg, sg2, isc = O.get_chi0_soc(st['Vmat'], st['Gk'], st['olist'], st['slist'], st['kmap'], st['invk'],
                                       st['invs'], st['temp'], st['Nx'], st['Ny'], st['Nz'])
    for it in (1, 2, 5):
        kw = dict(eps=1e-30, itemax=it, arnoldi_m=0)
        da = F.linearized_eliashberg_soc(*_eliash_args(st, chi.copy(), sg, sg2, isc, gs), **kw)
        la = F.last_lambda()
        db, lb = O.linearized_eliashberg_soc(*_eliash_args(st, chi.copy(), sg, sg2, isc, gs), **kw)
        if abs(lb) < 1e-12:   # the first triplet step can annihilate the initial gap: nothing to compare
            continue
        assert abs(la - lb) < 1e-10 * abs(lb), (it, la, lb)
        assert rel(da, db) < 1e-7, it


@pytest.mark.parametrize("gs,arnoldi_m,tol", [(1, 0, 1e-8), (1, 10, 1e-6), (-1, 10, 1e-6)])
def test_lin_eliash_soc_matches_oracle(F, gs, arnoldi_m, tol):
    """Converged solves on the one-orbital Rashba model.  The Arnoldi result
    of a 10-dimensional Krylov space after a shift many times the eigenvalue
    carries rounding at about 1e-7 relative, so the two FP64 evaluations agree
    to 1e-6 there; the power method converges to 1e-8."""
    st = _model("rashba_1orb")
    chi, sg, sg2, isc = O.get_chi0_soc(st['Vmat'], st['Gk'], st['olist'], st['slist'], st['kmap'], st['invk'],
                                       st['invs'], st['temp'], st['Nx'], st['Ny'], st['Nz'])
    kw = dict(eps=1e-8, itemax=300, arnoldi_m=arnoldi_m)
    da = F.linearized_eliashberg_soc(*_eliash_args(st, chi.copy(), sg, sg2, isc, gs), **kw)
    la = F.last_lambda()
    db, lb = O.linearized_eliashberg_soc(*_eliash_args(st, chi.copy(), sg, sg2, isc, gs), **kw)
    print(f"gap_sym {gs} arnoldi {arnoldi_m}: lambda {la:.10f} (oracle {lb:.10f})")
    assert abs(la - lb) < tol * abs(lb)
    assert rel(da, db) < 100 * tol


def test_soc_gap_helpers_match_oracle(F):
    st = _model("soc_2orb")
    rng = np.random.default_rng(3)
    No, Nw, Nkall = st['Norb'], st['Nw'], len(st['kmap'])
    d = rng.standard_normal((No, No, Nw, Nkall)) + 1j * rng.standard_normal((No, No, Nw, Nkall))
    sg, _ = O.soc_signs(st['olist'], st['slist'])
    for gs in (1, -1, -10):
        a = F.gen_Fk_soc(st['Gk'], d, st['invk'], st['invs'], st['slist'], gs)
        b = O.mkfk_trs_soc(st['Gk'], d, sg, st['slist'], st['invk'], st['invs'], gs)
        assert rel(a, b) < 1e-13
    a = F.conv_delta_orb_to_band_soc(d, st['uni'], st['invk'], st['invs'], st['slist'])
    b = O.conv_delta_orb_to_band_soc(d, st['uni'], st['invk'], st['invs'], st['slist'])
    assert rel(a, b) < 1e-13
