"""GPU parity of libgmod.so against the CPU oracle, through the C-ABI (ctypes).

FP64 path: every quantity must agree with the oracle to 1e-10 relative to its
largest element (the reference's own path is FP64; differences are FFT and LU
roundoff).  Shapes cover power-of-two and other FFT lengths, the four-step
split of long tau axes (2Nw > 256), 2D and 3D meshes, 1-4 orbitals."""
import numpy as np
import pytest

import oracle.pyrpa_oracle as O
from tests import models as MD

pytestmark = pytest.mark.gpu

TOL = 1e-10


@pytest.fixture(scope="module")
def F():
    import pyrpa_amd.flibs as F
    F.set_precision(0)
    return F


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


MODELS = {
    "1orb_4x4_nw16": lambda: MD.one_orbital(Nx=4, Ny=4, Nw=16, temp=0.1, mu=0.2),
    "1orb_8x8_nw160_fourstep": lambda: MD.one_orbital(Nx=8, Ny=8, Nw=160, temp=0.03, mu=0.3),
    "1orb_3x3_nw5_nonpow2": lambda: MD.one_orbital(Nx=3, Ny=3, Nw=5, temp=0.1, mu=0.1),
    "2orb_4x4_nw8": lambda: MD.two_orbital(Nx=4, Ny=4, Nw=8, temp=0.1, mu=0.0, U=0.5, J=0.1),
    "3orb_4x4x2_nw6": lambda: MD.random_model(3, 4, 4, 2, 6, 0.15, 0.05, 0.8, 0.15, seed=7),
    "4orb_2x2x4_nw4": lambda: MD.random_model(4, 2, 2, 4, 4, 0.2, 0.0, 0.6, 0.1, seed=8),
}


@pytest.mark.parametrize("name", list(MODELS))
def test_chi0_conv_and_stoner(F, name):
    st = MODELS[name]()
    args = (st['Gk'], st['kmap'], st['invk'], st['olist'], st['temp'], st['Nx'], st['Ny'], st['Nz'])
    a = F.get_chi0_conv(*args)
    b = O.get_chi0_conv(*args)
    assert rel(a, b) < TOL, name
    c, s1 = F.get_chi0(st['Smat'], st['Cmat'], st['Gk'], st['olist'], st['kmap'], st['invk'], st['temp'],
                       st['Nx'], st['Ny'], st['Nz'])
    assert rel(c, b) < TOL
    s2 = O.ckchi(b, st['Smat'], st['Cmat'])
    assert abs(s1 - s2) < 1e-10 * max(1, abs(s2))


@pytest.mark.parametrize("name", ["1orb_4x4_nw16", "2orb_4x4_nw8", "3orb_4x4x2_nw6"])
def test_chi0_tail(F, name):
    st = MODELS[name]()
    G = st['Gk'] * 1.0
    # an interacting-looking G: perturb G0 with a smooth self-energy so conv[G] != conv[G0]
    rng = np.random.default_rng(0)
    sig = 0.05 * (rng.standard_normal(G.shape) - 1j * np.abs(rng.standard_normal(G.shape)))
    sig = 0.5 * (sig + np.conj(np.swapaxes(sig, 0, 1)))
    G = O.gen_green(sig, st['hamk'], st['mu'], st['temp'])
    a, sa = F.get_chi0_tail(st['Smat'], st['Cmat'], G, st['eig'], st['uni'], st['olist'], st['kmap'], st['invk'],
                            st['mu'], st['temp'], st['Nx'], st['Ny'], st['Nz'])
    b, sb = O.get_chi0_tail(st['Smat'], st['Cmat'], G, st['eig'], st['uni'], st['olist'], st['kmap'], st['invk'],
                            st['mu'], st['temp'], st['Nx'], st['Ny'], st['Nz'])
    assert rel(a, b) < TOL
    assert abs(sa - sb) < 1e-10 * max(1, abs(sb))


def test_chi0_tail_exact_at_q0(F):
    """tests/test_rpa_flex.py:336-354 through the HIP library."""
    st = MD.one_orbital(Nx=8, Ny=8, Nw=32, temp=0.04, mu=0.3)
    cn, _ = F.get_chi0_tail(st['Smat'], st['Cmat'], st['Gk'], st['eig'], st['uni'], st['olist'], st['kmap'],
                            st['invk'], st['mu'], st['temp'], 8, 8, 1)
    ex0 = MD.lindhard_exact(st, [0])[0, 0]
    assert abs(cn[0, 0, 0, 0] - ex0) < 1e-10


@pytest.mark.parametrize("name", ["1orb_4x4_nw16", "2orb_4x4_nw8", "3orb_4x4x2_nw6", "4orb_2x2x4_nw4"])
def test_rpa_vertices_and_susceptibilities(F, name):
    st = MODELS[name]()
    chi = O.get_chi0_conv(st['Gk'], st['kmap'], st['invk'], st['olist'], st['temp'], st['Nx'], st['Ny'], st['Nz'])
    V = F.get_Vsigma_nosoc_flex(chi.copy(), st['Smat'], st['Cmat'])
    assert rel(V, O.get_Vsigma_nosoc_flex(chi, st['Smat'], st['Cmat'])) < TOL
    a = F.get_chis_chic(chi, st['Smat'], st['Cmat'])
    b = O.get_chis_chic(chi, st['Smat'], st['Cmat'])
    assert rel(a[0], b[0]) < TOL and rel(a[1], b[1]) < TOL
    chi0q = np.ascontiguousarray(np.transpose(chi[:, :, :, 3], (2, 0, 1)))
    assert rel(F.get_chis(chi0q, st['Smat']), O.get_chis(chi0q, st['Smat'])) < TOL
    for sw in (True, False):
        x = F.get_eig_or_tr_chi(a[0], st['invk'], sw)
        y = O.get_eig_or_tr_chi(b[0], st['invk'], sw)
        assert rel(x, y) < 1e-9


def test_scalar_rpa_chis_matches_closed_form(F):
    """tests/test_rpa_flex.py:119-127 through the HIP library."""
    chi0 = np.array([[[0.20 + 0.10j]], [[0.05 + 0.02j]]])
    chis = F.get_chis(chi0, np.array([[1.2]]))
    assert np.allclose(chis[:, 0, 0], chi0[:, 0, 0] / (1 - 1.2 * chi0[:, 0, 0]), rtol=1e-14)


def test_green_functions(F):
    st = MODELS["3orb_4x4x2_nw6"]()
    G0 = F.gen_Green0(st['eig'], st['uni'], 0.07, st['temp'], 6)
    assert rel(G0, O.gen_Green0(st['eig'], st['uni'], 0.07, st['temp'], 6)) < 1e-13
    rng = np.random.default_rng(1)
    sig = 0.1 * (rng.standard_normal(G0.shape) + 1j * rng.standard_normal(G0.shape))
    G = F.gen_green(sig, st['hamk'], 0.07, st['temp'])
    assert rel(G, O.gen_green(sig, st['hamk'], 0.07, st['temp'])) < 1e-12


@pytest.mark.parametrize("case", ["1orb_conv", "1orb_tail", "2orb_diis_sub2", "3orb_3d_tail"])
def test_mkself_matches_oracle(F, case):
    if case == "1orb_conv":
        st, U, J, fill, kw = MD.one_orbital(Nx=4, Ny=4, Nw=16, temp=0.1), 1.5, 0.0, 0.4, {}
    elif case == "1orb_tail":
        st, U, J, fill, kw = MD.one_orbital(Nx=4, Ny=4, Nw=16, temp=0.1), 1.0, 0.0, 0.4, dict(sw_tail=True)
    elif case == "2orb_diis_sub2":
        st, U, J, fill, kw = MD.two_orbital(Nx=4, Ny=4, Nw=8, temp=0.1), 0.8, 0.1, 0.9, dict(m_diis=3, sub_sigma=2)
    else:
        st = MD.random_model(3, 4, 4, 2, 6, 0.15, 0.0, 0, 0, seed=7)
        U, J, fill, kw = 0.6, 0.1, 1.2, dict(sw_tail=True, sw_rescale=True)
    S, C = O.get_scmat(st['olist'], st['site'], U, J)
    args = (S, C, st['kmap'], st['invk'], st['olist'], st['hamk'], st['eig'], st['uni'], 0.0, fill, st['temp'],
            st['Nw'], st['Nx'], st['Ny'], st['Nz'], False, False)
    opts = dict(eps=1e-10, pp=0.5, m_diis=1, scf_loop=200)
    opts.update(kw)
    sa, ma = F.mkself(*args, **opts)
    sb, mb, info = O.mkself(*args, return_info=True, **opts)
    stats = F.flex_stats()
    assert stats['iterations'] == info['iters']
    assert rel(sa, sb) < 1e-8
    assert abs(ma - mb) < 1e-8


def _flex_case(case):
    if case == "1orb_conv":
        st, U, J, fill, kw = MD.one_orbital(Nx=8, Ny=8, Nw=32, temp=0.05), 1.5, 0.0, 0.4, {}
    elif case == "1orb_tail":
        st, U, J, fill, kw = MD.one_orbital(Nx=8, Ny=8, Nw=32, temp=0.05), 1.5, 0.0, 0.4, dict(sw_tail=True)
    elif case == "2orb_diis_sub2":
        st, U, J, fill, kw = MD.two_orbital(Nx=8, Ny=8, Nw=16, temp=0.05), 0.8, 0.1, 0.9, dict(m_diis=3, sub_sigma=2)
    else:
        st = MD.random_model(3, 4, 4, 2, 16, 0.1, 0.0, 0, 0, seed=7)
        U, J, fill, kw = 0.6, 0.1, 1.2, dict(sw_tail=True, sw_rescale=True)
    S, C = O.get_scmat(st['olist'], st['site'], U, J)
    args = (S, C, st['kmap'], st['invk'], st['olist'], st['hamk'], st['eig'], st['uni'], 0.0, fill, st['temp'],
            st['Nw'], st['Nx'], st['Ny'], st['Nz'], False, False)
    return args, kw


@pytest.mark.parametrize("case", ["1orb_conv", "1orb_tail", "2orb_diis_sub2", "3orb_3d_tail"])
def test_mkself_mixed_precision_matches_fp64(F, case):
    """FP64 values at a reference iterate plus FP32 transforms of the
    differences from it: converges to the reference criterion esterr < 1e-8
    and lands on the FP64 fixed point (max|dSigma| / max|Sigma| < 1e-6,
    |dmu| < 1e-6)."""
    args, kw = _flex_case(case)
    opts = dict(eps=1e-8, pp=0.5, m_diis=1, scf_loop=300)
    opts.update(kw)
    s64, m64 = F.mkself(*args, **opts)
    n64 = F.flex_stats()['iterations']
    try:
        F.set_precision(1)
        s32, m32 = F.mkself(*args, **opts)
        st = F.flex_stats()
    finally:
        F.set_precision(0)
    assert st['esterr'] < 1e-8
    assert st['increment_iterations'] >= 1
    assert st['iterations'] <= n64 + 3
    assert rel(s32, s64) < 1e-6
    assert abs(m32 - m64) < 1e-6


@pytest.mark.parametrize("model", ["2orb", "3orb_random", "3orb_random_asym"])
@pytest.mark.parametrize("prec", [0, 1])
def test_mkself_accumulating_sigma_pass_matches_per_pair_passes(F, prec, model, monkeypatch):
    """The Sigma step's accumulating last pass (all vertex pairs of one output
    transformed in one persistent pass, their sum kept in registers) against
    the per-pair passes that accumulate through memory
    (GMOD_SIGMA_UNFUSED=1), on a 16x16x4 x 64-slot grid, where the last pass
    of the forward transform fills every tile register (the only grids that
    take the accumulating path).  Four iterations of FP64 FLEX and of the
    mixed loop (FP64 first iteration, then FP32 increments): the same
    arithmetic in the same order up to contraction, so 1e-12 (FP64) and 1e-6
    (FP32 increments) of max|Sigma|.  The increments run in four forms: the
    contraction form (default: one stored transform per transposed vertex
    pair, all outputs in one streaming kernel), the grouped accumulating
    pass (GMOD_SIGMA_ACCUM=1), the accumulating pass with three operands per
    source (GMOD_SIGMA_UNGROUPED=1) and the per-pair passes.  The fused
    forms transform one vertex pair of each transposed pair and read the
    other as the conjugate reversal plus the transformed nu_0 mismatch; the
    3-orbital model with complex hoppings (time reversal broken) has a
    vertex that is not Hermitian at nu_0, so there the correction carries
    weight.  With one off-diagonal element of S perturbed (S not symmetric)
    the static slot is not, so the FP64 step and the reference-vertex
    transforms stay unshared while the vertex differences (static slot
    zero) are shared: the contraction reads both kinds of entries."""
    if model == "2orb":
        st = MD.two_orbital(Nx=16, Ny=16, Nz=4, Nw=32, temp=0.05)
        U, J, fill = 0.8, 0.1, 0.9
    else:  # 3orb_random, 3orb_random_asym
        st = MD.random_model(3, 16, 16, 4, 32, 0.1, 0.0, 0, 0, seed=7)
        U, J, fill = 0.6, 0.1, 1.2
    S, C = O.get_scmat(st['olist'], st['site'], U, J)
    if model == "3orb_random_asym":
        S = S.copy()
        S[0, 1] += 0.02
    args = (S, C, st['kmap'], st['invk'], st['olist'], st['hamk'], st['eig'], st['uni'], 0.0, fill, st['temp'],
            st['Nw'], st['Nx'], st['Ny'], st['Nz'], False, False)
    opts = dict(eps=1e-14, pp=0.5, m_diis=1, scf_loop=4)
    try:
        F.set_precision(prec)
        monkeypatch.delenv("GMOD_SIGMA_UNFUSED", raising=False)
        monkeypatch.delenv("GMOD_SIGMA_UNGROUPED", raising=False)
        monkeypatch.delenv("GMOD_SIGMA_ACCUM", raising=False)
        sf, mf = F.mkself(*args, **opts)
        nf = F.flex_stats()
        if prec == 1:
            # the grouped accumulating pass instead of the contraction form
            monkeypatch.setenv("GMOD_SIGMA_ACCUM", "1")
            sa, ma = F.mkself(*args, **opts)
            monkeypatch.delenv("GMOD_SIGMA_ACCUM")
            # the increments' accumulating pass reading all three operands per
            # source instead of the grouped reference-vertex term
            monkeypatch.setenv("GMOD_SIGMA_UNGROUPED", "1")
            sg, mg = F.mkself(*args, **opts)
            monkeypatch.delenv("GMOD_SIGMA_UNGROUPED")
        monkeypatch.setenv("GMOD_SIGMA_UNFUSED", "1")
        su, mu_ = F.mkself(*args, **opts)
        nu = F.flex_stats()
    finally:
        F.set_precision(0)
    assert nf['iterations'] == nu['iterations'] == 4
    if prec == 1:
        assert nf['increment_iterations'] >= 2
    tol = 1e-12 if prec == 0 else 1e-6
    assert rel(sf, su) < tol
    assert abs(mf - mu_) < tol
    if prec == 1:
        assert rel(sg, su) < tol and abs(mg - mu_) < tol
        assert rel(sa, su) < tol and abs(ma - mu_) < tol


def test_mkself_warm_start_from_memory_and_sigma_bin(F, tmp_path, monkeypatch):
    st = MD.one_orbital(Nx=4, Ny=4, Nw=16, temp=0.1)
    S, C = O.get_scmat(st['olist'], st['site'], 1.2, 0.0)
    args = (S, C, st['kmap'], st['invk'], st['olist'], st['hamk'], st['eig'], st['uni'], 0.0, 0.4, st['temp'],
            16, 4, 4, 1)
    monkeypatch.chdir(tmp_path)
    s1, m1 = F.mkself(*args, True, False, eps=1e-10, pp=0.5, m_diis=1)
    assert (tmp_path / "sigma.bin").exists()
    # restart from the file: converged in one iteration, same answer
    s2, m2 = F.mkself(*args, False, True, eps=1e-8, pp=0.5, m_diis=1)
    assert F.flex_stats()['iterations'] <= 2
    assert rel(s2, s1) < 1e-8
    # in-memory seed scaled by 0.5 converges back to the same fixed point
    F.set_sigma_seed(s1, m1, m1)
    s3, m3 = F.mkself(*args, False, True, eps=1e-10, pp=0.5, m_diis=1, sigma_scale=0.5)
    assert rel(s3, s1) < 1e-7 and abs(m3 - m1) < 1e-7


@pytest.mark.parametrize("arnoldi_m", [0, 10])
def test_lin_eliash_matches_oracle(F, arnoldi_m):
    st = MD.one_orbital(Nx=4, Ny=4, Nw=8, temp=0.05, mu=-0.2, tx=1.0, ty=1.0, U=2.0)
    chi, _ = O.get_chi0(st['Smat'], st['Cmat'], st['Gk'], st['olist'], st['kmap'], st['invk'], st['temp'], 4, 4, 1)
    init = O.gap_symms(st['klist'], 1, 1)
    args = (st['Gk'], st['uni'], init, st['Smat'], st['Cmat'], st['olist'], st['plist'], st['kmap'], st['invk'],
            4, 4, 1, st['temp'], 1)
    ca = chi.copy()
    da, la = F.linearized_eliashberg(ca, *args, eps=1e-9, itemax=500, arnoldi_m=arnoldi_m)
    db, lb = O.linearized_eliashberg(chi.copy(), *args, eps=1e-9, itemax=500, arnoldi_m=arnoldi_m)
    assert abs(la - lb) < 1e-8 * max(1, abs(lb))
    # gap vectors agree up to a global sign (power method) / exactly (Arnoldi)
    s = np.sign(np.real(np.vdot(db.ravel(), da.ravel())))
    assert rel(da * s, db) < 1e-6
    assert rel(ca, O.get_V_delta(chi, st['Smat'], st['Cmat'])) < TOL  # chi leaves as V_Delta


@pytest.mark.parametrize("gap_sym", [1, -1])
def test_lin_eliash_multiorbital_arnoldi(F, gap_sym):
    st = MD.two_orbital(Nx=4, Ny=4, Nw=4, temp=0.1, mu=0.0, U=1.0, J=0.2)
    chi, _ = O.get_chi0(st['Smat'], st['Cmat'], st['Gk'], st['olist'], st['kmap'], st['invk'], st['temp'], 4, 4, 1)
    init = O.gap_symms(st['klist'], 2, gap_sym)
    args = (st['Gk'], st['uni'], init, st['Smat'], st['Cmat'], st['olist'], st['plist'], st['kmap'], st['invk'],
            4, 4, 1, st['temp'], gap_sym)
    ca = chi.copy()
    _, la = F.linearized_eliashberg(ca, *args, arnoldi_m=6)
    _, lb = O.linearized_eliashberg(chi.copy(), *args, arnoldi_m=6)
    assert abs(la - lb) < 1e-8 * max(1, abs(lb))
    assert rel(ca, O.get_V_delta(chi, st['Smat'], st['Cmat'], singlet=gap_sym >= 0)) < TOL


def test_lin_eliash_zero_interaction_has_zero_lambda(F):
    """tests/test_rpa_flex.py:187-205 through the HIP library."""
    st = MD.one_orbital(Nx=2, Ny=2, Nw=4)
    chi, _ = F.get_chi0(st['Smat'], st['Cmat'], st['Gk'], st['olist'], st['kmap'], st['invk'], st['temp'], 2, 2, 1)
    init = O.gap_symms(st['klist'], 1, 0)
    gap, lam = F.linearized_eliashberg(chi, st['Gk'], st['uni'], init, st['Smat'], st['Cmat'], st['olist'],
                                       st['plist'], st['kmap'], st['invk'], 2, 2, 1, st['temp'], 0, itemax=3,
                                       arnoldi_m=2)
    assert abs(lam) < 1e-12 and np.isfinite(gap).all()


def test_anomalous_and_initial_gap(F):
    st = MODELS["3orb_4x4x2_nw6"]()
    rng = np.random.default_rng(4)
    D = rng.standard_normal(st['Gk'].shape) + 1j * rng.standard_normal(st['Gk'].shape)
    assert rel(F.gen_Fk(st['Gk'], D, st['invk']), O.mkfk_trs_nsoc(st['Gk'], D)) < 1e-13
    init = rng.uniform(-1, 1, (3, len(st['klist'])))
    for gs in (0, 1, -1):
        a = F.get_initial_delta(init, st['uni'], st['kmap'], st['invk'], 6, gs)
        b = O.get_initial_delta(init, st['uni'], st['kmap'], st['invk'], 6, gs)
        assert rel(a, b) < 1e-13


def test_singular_denominator_raises_instead_of_killing_the_process(F):
    """The reference STOPs on zgesv failure (fchi.f90:347); here it raises."""
    chi0 = np.array([[[1.0 / 1.2 + 0j]]])
    with pytest.raises(RuntimeError):
        F.get_chis(chi0, np.array([[1.2]]))
    # the library keeps working afterwards
    assert np.isfinite(F.get_chis(np.array([[[0.1 + 0j]]]), np.array([[1.2]]))).all()
