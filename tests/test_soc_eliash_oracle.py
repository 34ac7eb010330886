"""The oracle's spinful linearized Eliashberg (feliashwithsoc.f90) pinned by
the mathematics.

  * Spin-degenerate singlet reduction: on a one-orbital Hubbard model with
    H(k) (x) 1_spin, the spinful gap operator applied to a singlet
    Delta_ud = -Delta_du = Delta equals the spinless singlet operator whose
    vertex is shifted by U (static slot 2U), at every k of the full zone:
    the spinful vertex -V + V chi_RPA V carries the bare U of both the
    density and the spin-flip channel.  The anomalous function, the -k
    unfolding with spin-flipped pairs and signs, the pair selection and the
    spin symmetrisation all enter it; the equal-spin blocks stay zero.
  * The power method, the two-pass Arnoldi and a dense eigen-decomposition
    of the same operator give one eigenvalue (singlet and triplet, with
    spin-orbit terms)."""
import numpy as np
import pytest

import oracle.pyrpa_oracle as O
from tests import models as MD


def _soc_chi(st):
    return O.get_chi0_soc(st['Vmat'], st['Gk'], st['olist'], st['slist'], st['kmap'], st['invk'], st['invs'],
                          st['temp'], st['Nx'], st['Ny'], st['Nz'])


def test_singlet_operator_reduces_to_the_spinless_one_with_a_shifted_vertex():
    Nx = Ny = 6
    Nw, U = 8, 1.5
    base = MD.one_orbital(Nx=Nx, Ny=Ny, Nw=Nw, temp=0.08, mu=0.0, ty=1.0, U=U)
    st = MD.spinful(base, U=U)
    kmap, invk = base['kmap'], base['invk']
    Nk, Nkall = len(base['klist']), len(kmap)
    rng = np.random.default_rng(0)
    dn = rng.standard_normal((1, 1, Nw, Nk)) + 1j * rng.standard_normal((1, 1, Nw, Nk))
    chi_s, sg, sg2, isc = _soc_chi(st)
    Vd_s = O.get_V_soc_flex(chi_s, st['Vmat'])
    full = dn[0, 0][:, invk[:, 0] - 1]                # even parity: Delta(-k) = Delta(k)
    ds = np.zeros((2, 2, Nw, Nkall), dtype=complex)
    ds[1, 0], ds[0, 1] = full, -full                  # Python [b, a]: Delta_ud, Delta_du
    fs = O.mkfk_trs_soc(st['Gk'], ds, sg, st['slist'], invk, st['invs'], 1)
    nd = np.asarray(O.mkdelta_soc(fs, Vd_s, st['Vmat'], kmap, invk, st['invs'], isc, st['olist'], st['slist'],
                                  Nx, Ny, 1, 1))
    chi_n = O.get_chi0_conv(base['Gk'], kmap, invk, base['olist'], base['temp'], Nx, Ny, 1)
    Vd_n = np.asarray(O.get_V_delta(chi_n, base['Smat'], base['Cmat'], True)).T + U
    fn = O.mkfk_trs_nsoc(base['Gk'], dn)
    ref = np.asarray(O.mkdelta_nsoc(fn, np.ascontiguousarray(Vd_n.T), 2 * base['Smat'], 2 * base['Cmat'], kmap,
                                    invk, np.ones(1), base['olist'], Nx, Ny, 1, True))[:, :, 0, 0].T
    scale = np.abs(ref).max()
    assert np.abs(nd[1, 0] + ref[:, invk[:, 0] - 1]).max() < 1e-12 * scale
    assert np.abs(nd[0, 1] - ref[:, invk[:, 0] - 1]).max() < 1e-12 * scale
    assert np.abs(nd[0, 0]).max() == 0.0 and np.abs(nd[1, 1]).max() == 0.0
    # without the shift the operators differ by the U term (the identity is not vacuous)
    ref0 = np.asarray(O.mkdelta_nsoc(fn, np.ascontiguousarray((Vd_n - U).T), base['Smat'], base['Cmat'], kmap, invk,
                                     np.ones(1), base['olist'], Nx, Ny, 1, True))[:, :, 0, 0].T
    assert np.abs(nd[1, 0] + ref0[:, invk[:, 0] - 1]).max() > 1e-3 * scale


def _soc_model():
    base = MD.one_orbital(Nx=4, Ny=4, Nw=4, temp=0.1, mu=0.1)
    return MD.spinful(base, alpha=0.25, U=2.0)


def _args(st, chi, sg, sg2, isc, gs):
    init = O.gap_symms(st['klist'], st['Norb'], gs)
    return (chi, st['Gk'], st['uni'], init, st['Vmat'], sg, sg2, np.ones(st['Norb']), st['slist'], st['olist'],
            st['kmap'], st['invk'], st['invs'], isc, st['Nx'], st['Ny'], st['Nz'], st['temp'], gs)


@pytest.mark.parametrize("gs", [1, -1])
def test_power_arnoldi_and_dense_operator_agree(gs):
    st = _soc_model()
    chi, sg, sg2, isc = _soc_chi(st)
    a = _args(st, chi, sg, sg2, isc, gs)
    _, lam_a = O.linearized_eliashberg_soc(*a, arnoldi_m=30)
    _, lam_p = O.linearized_eliashberg_soc(*a, eps=1e-10, itemax=3000, arnoldi_m=0)
    # the dense operator on the full zone; the spin symmetrisation conjugates
    # components, so the map is real-linear: its matrix on (Re, Im) coordinates
    Vd = O.get_V_soc_flex(chi, st['Vmat'])
    No, Nw, Nkall = st['Norb'], st['Nw'], len(st['kmap'])
    n = No * No * Nw * Nkall
    A = np.zeros((2 * n, 2 * n))
    w = st['temp'] / Nkall
    for c in range(2 * n):
        e = np.zeros(n, dtype=complex)
        e[c % n] = 1.0 if c < n else 1.0j
        fk = O.mkfk_trs_soc(st['Gk'], e.reshape(No, No, Nw, Nkall), sg, st['slist'], st['invk'], st['invs'], gs)
        y = -w * np.asarray(O.mkdelta_soc(fk, Vd, st['Vmat'], st['kmap'], st['invk'], st['invs'], isc,
                                          st['olist'], st['slist'], st['Nx'], st['Ny'], st['Nz'], gs)).ravel()
        A[:n, c], A[n:, c] = y.real, y.imag
    ev = np.linalg.eigvals(A)
    lam_d = ev.real.max()
    print(f"gap_sym {gs}: Arnoldi {lam_a:.10f}, power {lam_p:.10f}, dense {lam_d:.10f}")
    # both solvers reach the leading eigenvalue of the singlet operator; for
    # this triplet one the shifted power method stops short of it (its exit
    # test for small positive eigenvalues, feliashwithsoc.f90:82-93, needs
    # thirty near-exact steps) while the Arnoldi solver reaches it.  The
    # solvers are the reference's: the parity tests reproduce, not correct, them.
    assert abs(lam_a - lam_d) < 1e-6 * max(1.0, abs(lam_d))   # a 30-dimensional Krylov space
    if gs > 0:
        assert abs(lam_p - lam_d) < 1e-6 * max(1.0, abs(lam_d))
