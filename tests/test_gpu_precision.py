"""GPU parity of the FP32-convolution paths and of the 5-orbital (Nchi = 25)
path against the FP64 oracle, through the C-ABI.

FP32 transforms carry a relative rounding error of a few 1e-7 per pass, so
the FP32 chi0 is checked to TOL32 = 2e-5 of max|chi0|, its Stoner factor to
1e-4 relative; the tail-corrected chi0 keeps its analytic tau reference and
sawtooth term in FP64.  lambda of the linearized Eliashberg equation with
FP32 convolutions (FP64 eigen-solver, norms and inner products) is checked
to 1e-5 absolute, the accuracy the FP64 result is asked to have.  The
5-orbital vertex, susceptibilities and FLEX iterations run in FP64 and are
checked to 1e-10 like the other FP64 paths."""
import numpy as np
import pytest

import oracle.pyrpa_oracle as O
from tests import models as MD

pytestmark = pytest.mark.gpu

TOL32 = 2e-5
TOL64 = 1e-10


@pytest.fixture(scope="module")
def F():
    import pyrpa_amd.flibs as F
    F.set_precision(0)
    yield F
    F.set_precision(0)


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


MODELS = {
    "1orb_8x8_nw32": lambda: MD.one_orbital(Nx=8, Ny=8, Nw=32, temp=0.05, mu=0.3),
    "2orb_4x4_nw8": lambda: MD.two_orbital(Nx=4, Ny=4, Nw=8, temp=0.1, mu=0.0, U=0.5, J=0.1),
    "3orb_4x4x2_nw6": lambda: MD.random_model(3, 4, 4, 2, 6, 0.15, 0.05, 0.8, 0.15, seed=7),
    "1orb_8x8_nw160_fourstep": lambda: MD.one_orbital(Nx=8, Ny=8, Nw=160, temp=0.03, mu=0.3),
}


def _interacting_green(st, seed=0, amp=0.05):
    rng = np.random.default_rng(seed)
    G = st['Gk']
    sig = amp * (rng.standard_normal(G.shape) - 1j * np.abs(rng.standard_normal(G.shape)))
    sig = 0.5 * (sig + np.conj(np.swapaxes(sig, 0, 1)))
    return O.gen_green(sig, st['hamk'], st['mu'], st['temp'])


@pytest.mark.parametrize("name", list(MODELS))
def test_fp32_chi0_conv_and_stoner(F, name):
    st = MODELS[name]()
    ref = O.get_chi0_conv(st['Gk'], st['kmap'], st['invk'], st['olist'], st['temp'], st['Nx'], st['Ny'], st['Nz'])
    sref = O.ckchi(ref, st['Smat'], st['Cmat'])
    try:
        F.set_precision(1)
        a = F.get_chi0_conv(st['Gk'], st['kmap'], st['invk'], st['olist'], st['temp'], st['Nx'], st['Ny'], st['Nz'])
        c, s1 = F.get_chi0(st['Smat'], st['Cmat'], st['Gk'], st['olist'], st['kmap'], st['invk'], st['temp'],
                           st['Nx'], st['Ny'], st['Nz'])
    finally:
        F.set_precision(0)
    assert rel(a, ref) < TOL32, rel(a, ref)
    assert rel(c, ref) < TOL32
    assert abs(s1 - sref) < 1e-4 * max(1.0, abs(sref))
    # the FP32 result is not the FP64 one: the precision switch took effect
    assert rel(a, ref) > 1e-13


@pytest.mark.parametrize("name", ["1orb_8x8_nw32", "2orb_4x4_nw8", "3orb_4x4x2_nw6"])
def test_fp32_chi0_tail(F, name):
    st = MODELS[name]()
    G = _interacting_green(st)
    args = (st['Smat'], st['Cmat'], G, st['eig'], st['uni'], st['olist'], st['kmap'], st['invk'], st['mu'],
            st['temp'], st['Nx'], st['Ny'], st['Nz'])
    b, sb = O.get_chi0_tail(*args)
    try:
        F.set_precision(1)
        a, sa = F.get_chi0_tail(*args)
    finally:
        F.set_precision(0)
    assert rel(a, b) < TOL32, rel(a, b)
    assert abs(sa - sb) < 1e-4 * max(1.0, abs(sb))


def test_fp32_chi0_tail_exact_at_q0(F):
    """The tail-corrected chi0 at q = 0, nu = 0 equals the exact Lindhard
    value (tests/test_rpa_flex.py:336-354): with FP32 transforms to FP32
    accuracy, the analytic reference being FP64."""
    st = MD.one_orbital(Nx=8, Ny=8, Nw=32, temp=0.04, mu=0.3)
    try:
        F.set_precision(1)
        cn, _ = F.get_chi0_tail(st['Smat'], st['Cmat'], st['Gk'], st['eig'], st['uni'], st['olist'], st['kmap'],
                                st['invk'], st['mu'], st['temp'], 8, 8, 1)
    finally:
        F.set_precision(0)
    ex0 = MD.lindhard_exact(st, [0])[0, 0]
    assert abs(cn[0, 0, 0, 0] - ex0) < 1e-6 * abs(ex0)


def _eliash_case(kind):
    if kind == "1orb":
        st = MD.one_orbital(Nx=4, Ny=4, Nw=8, temp=0.05, mu=-0.2, tx=1.0, ty=1.0, U=2.0)
        gs = 1
    else:
        st = MD.two_orbital(Nx=4, Ny=4, Nw=4, temp=0.1, mu=0.0, U=1.0, J=0.2)
        gs = 1
    chi, _ = O.get_chi0(st['Smat'], st['Cmat'], st['Gk'], st['olist'], st['kmap'], st['invk'], st['temp'],
                        st['Nx'], st['Ny'], st['Nz'])
    init = O.gap_symms(st['klist'], st['Norb'], gs)
    args = (st['Gk'], st['uni'], init, st['Smat'], st['Cmat'], st['olist'], st['plist'], st['kmap'], st['invk'],
            st['Nx'], st['Ny'], st['Nz'], st['temp'], gs)
    return chi, args


@pytest.mark.parametrize("kind", ["1orb", "2orb"])
@pytest.mark.parametrize("arnoldi_m", [0, 10])
def test_fp32_lin_eliash_lambda(F, kind, arnoldi_m):
    chi, args = _eliash_case(kind)
    m = arnoldi_m if kind == "1orb" else (6 if arnoldi_m else 0)
    _, lb = O.linearized_eliashberg(chi.copy(), *args, eps=1e-9, itemax=500, arnoldi_m=m)
    try:
        F.set_precision(1)
        ca = chi.copy()
        da, la = F.linearized_eliashberg(ca, *args, eps=1e-9, itemax=500, arnoldi_m=m)
    finally:
        F.set_precision(0)
    assert abs(la - lb) < 1e-5, (la, lb)
    assert np.isfinite(da).all()
    # the vertex and the eigen-solver are FP64: chi still leaves as V_Delta to FP64 accuracy
    assert rel(ca, O.get_V_delta(chi, args[3], args[4])) < TOL64


def test_fp32_mkself_full_fields(F):
    """precision 2: FP32 transforms of the full fields (no FP64 accumulators);
    converges to the FP32 floor and lands near the FP64 fixed point."""
    st = MD.one_orbital(Nx=8, Ny=8, Nw=32, temp=0.05)
    S, C = O.get_scmat(st['olist'], st['site'], 1.5, 0.0)
    args = (S, C, st['kmap'], st['invk'], st['olist'], st['hamk'], st['eig'], st['uni'], 0.0, 0.4, st['temp'],
            st['Nw'], st['Nx'], st['Ny'], st['Nz'], False, False)
    opts = dict(eps=1e-5, pp=0.5, m_diis=1, scf_loop=100, sw_tail=True)
    s64, m64 = F.mkself(*args, **opts)
    try:
        F.set_precision(2)
        s32, m32 = F.mkself(*args, **opts)
        stats = F.flex_stats()
    finally:
        F.set_precision(0)
    assert stats['esterr'] < 1e-5
    assert rel(s32, s64) < 1e-4 and abs(m32 - m64) < 1e-4


# ------------------------------------------------------------------ 5 orbitals (Nchi = 25)
@pytest.fixture(scope="module")
def five():
    return MD.random_model(5, 4, 4, 1, 4, 0.1, 0.0, 0.6, 0.15, seed=11, amp=0.25)


def test_five_orbital_vertex_and_susceptibilities(F, five):
    st = five
    chi = O.get_chi0_conv(st['Gk'], st['kmap'], st['invk'], st['olist'], st['temp'], 4, 4, 1)
    assert chi.shape[0] == 25
    V = F.get_Vsigma_nosoc_flex(chi.copy(), st['Smat'], st['Cmat'])
    assert rel(V, O.get_Vsigma_nosoc_flex(chi, st['Smat'], st['Cmat'])) < TOL64
    a = F.get_chis_chic(chi, st['Smat'], st['Cmat'])
    b = O.get_chis_chic(chi, st['Smat'], st['Cmat'])
    assert rel(a[0], b[0]) < TOL64 and rel(a[1], b[1]) < TOL64
    xs, xc = F.get_chis_chic_all(chi, st['Smat'], st['Cmat'])
    ys, yc = O.get_chis_chic_all(chi, st['Smat'], st['Cmat'])
    assert rel(xs, ys) < TOL64 and rel(xc, yc) < TOL64
    # nu = 0 slice of the all-frequency export is the static export
    assert rel(xs[:, :, 0, :], a[0]) < TOL64
    cc, s1 = F.get_chi0(st['Smat'], st['Cmat'], st['Gk'], st['olist'], st['kmap'], st['invk'], st['temp'], 4, 4, 1)
    assert rel(cc, chi) < TOL64
    assert abs(s1 - O.ckchi(chi, st['Smat'], st['Cmat'])) < 1e-10


def test_five_orbital_flex_iterations(F, five):
    """Three FLEX iterations of the 5-orbital model against the oracle's."""
    st = five
    S, C = O.get_scmat(st['olist'], st['site'], 0.6, 0.15)
    args = (S, C, st['kmap'], st['invk'], st['olist'], st['hamk'], st['eig'], st['uni'], 0.0, 2.4, st['temp'],
            st['Nw'], 4, 4, 1, False, False)
    opts = dict(eps=0.0, pp=0.5, m_diis=1, scf_loop=3)
    sa, ma = F.mkself(*args, **opts)
    sb, mb, info = O.mkself(*args, return_info=True, **opts)
    assert F.flex_stats()['iterations'] == info['iters'] == 3
    assert rel(sa, sb) < 1e-8
    assert abs(ma - mb) < 1e-8


def test_chis_all_matches_per_q_exports(F):
    """gmod_get_chis_chic_all agrees with get_chis at each q (all nu) and
    with get_chis_chic at nu = 0, the reference's own two entry points."""
    st = MD.two_orbital(Nx=4, Ny=4, Nw=6, temp=0.1, mu=0.0, U=1.0, J=0.2)
    chi = O.get_chi0_conv(st['Gk'], st['kmap'], st['invk'], st['olist'], st['temp'], 4, 4, 1)
    xs, xc = F.get_chis_chic_all(chi, st['Smat'], st['Cmat'])
    for k in range(chi.shape[3]):
        q = np.ascontiguousarray(np.transpose(chi[:, :, :, k], (2, 0, 1)))
        e = F.get_chis(q, st['Smat'])
        assert rel(np.transpose(xs[:, :, :, k], (2, 0, 1)), e) < TOL64
    s0, c0 = F.get_chis_chic(chi, st['Smat'], st['Cmat'])
    assert rel(xs[:, :, 0, :], s0) < TOL64 and rel(xc[:, :, 0, :], c0) < TOL64
    only_s, none_c = F.get_chis_chic_all(chi, st['Smat'], st['Cmat'], charge=False)
    assert none_c is None and rel(only_s, xs) == 0.0
