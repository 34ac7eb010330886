"""Wannier-R output of Sigma / Delta (libs/plibs/_wannier_io.py:48-216).

The oracle's transform is pinned by a closed form: data(k) = sum_R c_R
e^{-2 pi i k.R} transforms back to c_R at R (mod the grid) and zero elsewhere;
on irreducible points with time-reversal unfolding this holds for real c_R
(then data(-k) = conj data(k)), on full-grid data for any c_R.  The host
side (zero cut, R selection, .dat/.npz formats) is checked on the CPU; the
device transform (gmod_wannier_r) against the oracle on the GPU, FP64 in a
different summation order, to 1e-13 of the largest element."""
import numpy as np
import pytest

import oracle.pyrpa_oracle as O

MESHES = [(8, 8, 1), (6, 5, 2), (4, 4, 4), (5, 3, 1)]


def _closed_form(Nx, Ny, Nz, Nw, No, real, seed, klist_full):
    rng = np.random.default_rng(seed)
    Rs = [(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (1, -1, 0), (0, 0, 1)]
    c = {}
    for R in Rs:
        v = rng.standard_normal((No, No, Nw))
        if not real:
            v = v + 1j * rng.standard_normal((No, No, Nw))
        c[R] = v
    ph = lambda R: np.exp(-2j * np.pi * (klist_full @ np.array(R, dtype=float)))
    data = sum(c[R][..., None] * ph(R)[None, None, None, :] for R in Rs)
    return c, data


def _expected(c, Nx, Ny, Nz, No, Ncut):
    out = np.zeros((No, No, Ncut, Nx, Ny, Nz), dtype=complex)
    for R, v in c.items():
        out[:, :, :, R[0] % Nx, R[1] % Ny, R[2] % Nz] += v[:, :, :Ncut]
    return out


@pytest.mark.parametrize("mesh", MESHES)
def test_oracle_transform_closed_form(mesh):
    Nx, Ny, Nz = mesh
    klist, kmap, invk = O.gen_irr_k_TRS(Nx, Ny, Nz)
    kfull = kmap / np.array([Nx, Ny, Nz], dtype=float)
    No, Nw, Ncut = 2, 5, 3
    # irreducible points + time reversal: real coefficients
    c, data_full = _closed_form(Nx, Ny, Nz, Nw, No, True, 1, klist)
    got = O.wannier_r(data_full, kmap, invk, Nx, Ny, Nz, Ncut)
    assert np.abs(got - _expected(c, Nx, Ny, Nz, No, Ncut)).max() < 1e-12
    # full grid in kmap order: any coefficients
    c, data_full = _closed_form(Nx, Ny, Nz, Nw, No, False, 2, kfull)
    got = O.wannier_r(data_full, kmap, invk, Nx, Ny, Nz, 99, full=True)
    assert got.shape[2] == Nw
    assert np.abs(got - _expected(c, Nx, Ny, Nz, No, Nw)).max() < 1e-12


def test_select_cuts_and_centres():
    Nx, Ny, Nz = 6, 5, 2
    d = np.zeros((1, 1, 2, Nx, Ny, Nz), dtype=complex)
    d[0, 0, 0, 0, 0, 0] = 1.0 + 1e-7j          # imaginary part below 1e-5 * max: cut
    d[0, 0, 1, 4, 1, 1] = 0.5j                 # R = (-2, 1, 1): z = 1 = Nz/2 stays positive
    d[0, 0, 0, 3, 3, 0] = 1e-6                 # cut entirely: R dropped
    d[0, 0, 1, 3, 2, 0] = -2e-5                # kept: R = (3, 2, 0)
    out, rv = O.wannier_select(d, Nx, Ny, Nz, 1e-5)
    assert rv.tolist() == [[0, 0, 0], [3, 2, 0], [-2, 1, 1]]
    assert out[0, 0, 0, 0] == 1.0 and out[0, 0, 1, 2] == 0.5j and out[0, 0, 1, 1] == -2e-5
    from pyrpa_amd.wannier_io import _select_nonzero
    out2, rv2 = _select_nonzero(d.copy(), Nx, Ny, Nz, 1e-5)
    assert np.array_equal(rv, rv2) and np.array_equal(out, out2)


def test_dat_layout(tmp_path):
    from pyrpa_amd.wannier_io import _write_dat
    rng = np.random.default_rng(4)
    No, Ncut, Nr = 2, 3, 4
    d = rng.standard_normal((No, No, Ncut, Nr)) + 1j * rng.standard_normal((No, No, Ncut, Nr))
    rv = np.array([[0, 0, 0], [1, -1, 0], [-3, 2, 1], [12, -10, 3]], dtype=np.int64)
    f = str(tmp_path / "x")
    _write_dat(f, d, rv, 40, 'Self-energy', mu=0.125, temp=0.01)
    lines = open(f + ".dat").read().splitlines()
    assert lines[0] == '# Self-energy: Norb=2 N_iw=3 Nr=4 mu=0.12500000 temp=1.00000000e-02'
    assert lines[1] == '2 3 4'
    assert len(lines) == 2 + Nr * Ncut * No * No
    # line of (R index 2, frequency 1, orbitals (1, 0)): (ir, n, i, j) C order
    ln = 2 + ((2 * Ncut + 1) * No + 1) * No + 0
    v = d[1, 0, 1, 2]
    assert lines[ln] == '  -3   2   1  2  1    2' + '%16.8e%16.8e' % (v.real, v.imag)
    back = np.loadtxt(f + ".dat", skiprows=2)
    assert back.shape == (Nr * Ncut * No * No, 8)
    rec = back[:, 6] + 1j * back[:, 7]
    assert np.abs(rec - d.transpose(3, 2, 0, 1).ravel()).max() < 1e-7 * np.abs(d).max()


@pytest.mark.gpu
@pytest.mark.parametrize("mesh", MESHES)
def test_device_transform_matches_oracle(mesh):
    import pyrpa_amd.flibs as F
    Nx, Ny, Nz = mesh
    klist, kmap, invk = O.gen_irr_k_TRS(Nx, Ny, Nz)
    Nk, Nkall = len(klist), Nx * Ny * Nz
    rng = np.random.default_rng(sum(mesh))
    for No, Nw, Ncut in ((1, 8, 8), (3, 6, 4)):
        data = rng.standard_normal((No, No, Nw, Nk)) + 1j * rng.standard_normal((No, No, Nw, Nk))
        a = F.wannier_r(data, kmap, invk, Nx, Ny, Nz, Ncut)
        b = O.wannier_r(data, kmap, invk, Nx, Ny, Nz, Ncut)
        assert a.shape == b.shape
        assert np.abs(a - b).max() < 1e-13 * np.abs(b).max(), (mesh, No)
        full = rng.standard_normal((No, No, Nw, Nkall)) + 1j * rng.standard_normal((No, No, Nw, Nkall))
        a = F.wannier_r(full, kmap, invk, Nx, Ny, Nz, Ncut, full=True)
        b = O.wannier_r(full, kmap, invk, Nx, Ny, Nz, Ncut, full=True)
        assert np.abs(a - b).max() < 1e-13 * np.abs(b).max(), (mesh, No, "full")
    # the closed form through the device
    c, data = _closed_form(Nx, Ny, Nz, 4, 2, True, 7, klist)
    a = F.wannier_r(data, kmap, invk, Nx, Ny, Nz, 4)
    assert np.abs(a - _expected(c, Nx, Ny, Nz, 2, 4)).max() < 1e-12


@pytest.mark.gpu
def test_writers_match_oracle(tmp_path):
    from pyrpa_amd.wannier_io import output_gap_wannier, output_self_wannier
    Nx, Ny, Nz, Nw, No, temp = 8, 6, 1, 12, 2, 0.02
    klist, kmap, invk = O.gen_irr_k_TRS(Nx, Ny, Nz)
    rng = np.random.default_rng(12)
    # decaying in R, so the zero cut removes the far R vectors
    c, sig = _closed_form(Nx, Ny, Nz, Nw, No, True, 3, klist)
    sig = sig + 1e-9 * rng.standard_normal(sig.shape)
    f = str(tmp_path / "self")
    output_self_wannier(sig, 0.3, kmap, invk, Nx, Ny, Nz, Nw, temp, N_cut=8, zero_tol=1e-5, fname=f)
    z = np.load(f + ".npz")
    ref, rv = O.wannier_select(O.wannier_r(sig, kmap, invk, Nx, Ny, Nz, 8), Nx, Ny, Nz, 1e-5)
    # six R vectors, (0,0,1) folds onto (0,0,0) on a single layer: five survive
    assert np.array_equal(z["rvec"], rv) and len(rv) == 5
    assert np.abs(z["sigma"] - ref).max() < 1e-13 * np.abs(ref).max()
    assert np.allclose(z["iw"], (2 * np.arange(8) + 1) * np.pi * temp) and z["mu"] == 0.3 and z["temp"] == temp
    assert open(f + ".dat").readline().startswith('# Self-energy: Norb=2 N_iw=8 Nr=5 mu=0.30000000')
    # gap on every grid point (the spin-orbit layout)
    kfull = kmap / np.array([Nx, Ny, Nz], dtype=float)
    c, gap = _closed_form(Nx, Ny, Nz, Nw, No, False, 5, kfull)
    g = str(tmp_path / "gap")
    output_gap_wannier(gap, kmap, invk, Nx, Ny, Nz, Nw, temp, N_cut=64, fname=g)
    z = np.load(g + ".npz")
    ref, rv = O.wannier_select(O.wannier_r(gap, kmap, invk, Nx, Ny, Nz, 64, full=True), Nx, Ny, Nz, 1e-5)
    assert np.array_equal(z["rvec"], rv) and z["gap"].shape == (No, No, Nw, 5)
    assert np.abs(z["gap"] - ref).max() < 1e-13 * np.abs(ref).max()
