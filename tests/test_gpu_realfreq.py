"""GPU parity of the real-frequency susceptibility (get_chi_irr, chiq_map;
libs/src/fchi.f90:94-317) and of the Pade continuation (get_a_,
pade_with_trace, pade_analytic_continuation_arrays_; libs/src/fpade.f90)
against the oracle, through the C-ABI.  FP64 arithmetic on both sides in a
different summation order: chi0 to 1e-11 of its largest element.  The Pade
recursion is ill-conditioned (a relative change of 1e-15 in the data moves
exactly determined coefficients by about 1e-7 and the continuation near poles
0.05 from the axis by about 1e-6), so the coefficients agree to 1e-6 and the
continuations to 1e-4 of their largest value; a slip in the recursion gives
errors of order one."""
import numpy as np
import pytest

import oracle.pyrpa_oracle as O
from tests import models as MD
from tests.test_realfreq_oracle import full_mesh

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def F():
    import pyrpa_amd.flibs as F
    F.set_precision(0)
    return F


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


MODELS = {
    "1orb_8x8": lambda: MD.one_orbital(Nx=8, Ny=8, temp=0.05, mu=0.1),
    "2orb_6x6": lambda: MD.two_orbital(Nx=6, Ny=6, temp=0.1, mu=0.1, V=0.3, U=0.6, J=0.1),
    "3orb_4x4x2": lambda: MD.random_model(3, 4, 4, 2, 4, 0.1, 0.05, 0.5, 0.1, seed=5, real=True),
    "5orb_4x4": lambda: MD.random_model(5, 4, 4, 1, 4, 0.1, 0.0, 0.4, 0.1, seed=11, amp=0.25, real=True),
}


@pytest.mark.parametrize("name", list(MODELS))
def test_chi_irr_matches_oracle(F, name):
    st = MODELS[name]()
    kl, eig, uni, ff = full_mesh(st)
    wl = np.linspace(-1.5, 1.5, 13)
    for q in (st['klist'][len(st['klist']) // 3], [0.0, 0.0, 0.0]):
        qs = O.get_qshift(kl, q)
        a = F.get_chi_irr(uni, eig, ff, qs, st['olist'], wl, 0.05, st['temp'])
        b = O.get_chi_irr(uni, eig, ff, qs, st['olist'], wl, 0.05, st['temp'])
        assert rel(a, b) < 1e-11, (name, q)
    # the static point w = 0 takes the -f' branch for equal energies
    a = F.get_chi_irr(uni, eig, ff, O.get_qshift(kl, [0, 0, 0]), st['olist'], np.array([0.0]), 0.05, st['temp'])
    b = O.get_chi_irr(uni, eig, ff, O.get_qshift(kl, [0, 0, 0]), st['olist'], np.array([0.0]), 0.05, st['temp'])
    assert rel(a, b) < 1e-11


@pytest.mark.parametrize("name", ["1orb_8x8", "2orb_6x6"])
def test_chis_qmap_matches_oracle(F, name):
    st = MODELS[name]()
    kl, eig, uni, ff = full_mesh(st)
    args = (uni, eig, ff, kl, st['Smat'], st['olist'], st['Nx'], st['Ny'], st['temp'], 0.3, 0.05)
    a = F.chis_qmap(*args)
    b = O.chis_qmap(*args)
    assert rel(a[1], b[1]) < 1e-11 and rel(a[0], b[0]) < 1e-10


def _poles(M, seed):
    rng = np.random.default_rng(seed)
    poles = rng.uniform(-2.0, 2.0, M) - 1j * rng.uniform(0.05, 0.3, M)
    w = rng.uniform(0.2, 1.0, M)
    return lambda z: (w[None, :] / (np.asarray(z)[:, None] - poles[None, :])).sum(axis=1)


def test_pade_matches_oracle(F):
    T = 0.05
    wl = np.linspace(-3, 3, 97) + 0.05j
    R = _poles(6, 1)
    # coefficients of 12 points of a 6-pole function (the continued fraction
    # is exactly determined): a relative perturbation of 1e-15 of the data
    # moves them by about 1e-7, so two FP64 evaluations agree to 1e-6
    iw12 = 1j * (2 * np.arange(12) + 1) * np.pi * T
    a = F.get_a(R(iw12), iw12)
    assert rel(a, O.get_a(R(iw12), iw12)) < 1e-6
    # with more points than poles the trailing coefficients are round-off
    # (the same perturbation changes them by O(1)); only the continuation is
    # meaningful, and it reproduces the function
    iw = 1j * (2 * np.arange(16) + 1) * np.pi * T
    a = F.get_a(R(iw), iw)
    Q, P = O.get_QP(a, iw, wl)
    assert rel(P / Q, R(wl)) < 1e-4
    # orbital-traced continuation of a (2, 2, Niw, Nk) array, different data per k
    Nk = 5
    A = np.zeros((2, 2, len(iw), Nk), dtype=complex)
    for k in range(Nk):
        A[0, 0, :, k] = _poles(4, 10 + k)(iw)
        A[1, 1, :, k] = _poles(3, 20 + k)(iw)
        A[1, 0, :, k] = 3.0
    x = F.pade_with_trace(A, iw, wl)
    y = O.pade_with_trace(A, iw, wl)
    assert rel(x, y) < 1e-4
    # every row of an (Nk, Niw) array
    rows = A[0, 0]                       # Python [Niw, Nk] = Fortran arrayin(Nk, Niw)
    z = F.pade_arrays(rows, iw, wl)      # [Nw, Nk]
    for k in range(Nk):
        Q, P = O.get_QP(O.get_a(rows[:, k], iw), iw, wl)
        assert rel(z[:, k], P / Q) < 1e-4
