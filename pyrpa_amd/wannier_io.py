"""Wannier-R writers for Sigma(k, i w_n) and Delta(k, i w_n), mirroring the
reference's libs/plibs/_wannier_io.py:145-216 (same names, arguments, files
and printed lines).

The k -> R transform (unfolding the irreducible points with time reversal and
the normalised inverse FFT over the k grid) runs on the GPU through
``flibs.wannier_r``; the zero-tolerance cut, the R-vector selection and the
.npz / .dat formats are host work on the transformed array."""
import numpy as np

from . import flibs


def _select_nonzero(data_R, Nx, Ny, Nz, zero_tol):
    """Cut Re/Im parts below zero_tol * max|data_R| to zero and keep the grid
    points where any (orbital pair, frequency) survives, in C order of
    (x, y, z), with indices above N/2 wrapped to negative R
    (_wannier_io.py:84-99).  data_R is modified in place.  Returns
    (data_out[Norb, Norb, N_cut, Nr], rvec[Nr, 3] int64)."""
    cut = np.abs(data_R).max() * zero_tol
    re, im = data_R.real, data_R.imag
    re[np.abs(re) < cut] = 0.0
    im[np.abs(im) < cut] = 0.0
    keep = (data_R != 0).any(axis=(0, 1, 2))
    ix, iy, iz = np.nonzero(keep)
    rvec = np.stack([np.where(ix > Nx // 2, ix - Nx, ix), np.where(iy > Ny // 2, iy - Ny, iy),
                     np.where(iz > Nz // 2, iz - Nz, iz)], axis=1).astype(np.int64)
    return data_R[:, :, :, ix, iy, iz], rvec


def _write_dat(fname, data_out, rvec, Nw_orig, label, mu=None, temp=None):
    """Text file fname.dat in the extended ham_r.dat layout of
    _wannier_io.py:101-143: a comment header, 'Norb N_iw Nr', then one line
    'Rx Ry Rz io jo iw Re Im' per (R, frequency, orbital pair), R slowest and
    the second orbital fastest."""
    Norb, _, N_cut, Nr = data_out.shape
    head = f'# {label}: Norb={Norb} N_iw={N_cut} Nr={Nr}'
    if mu is not None:
        head += f' mu={mu:.8f}'
    if temp is not None:
        head += f' temp={temp:.8e}'
    # rows in (R, n, i, j) order
    vals = np.ascontiguousarray(data_out.transpose(3, 2, 0, 1)).ravel()
    r, n, i, j = (a.ravel() for a in np.indices((Nr, N_cut, Norb, Norb)))
    ints = np.stack([rvec[r, 0], rvec[r, 1], rvec[r, 2], i + 1, j + 1, n + 1], axis=1)
    body = [f'{a:4d}{b:4d}{c:4d}{o:3d}{p:3d}{w:5d}{x.real:16.8e}{x.imag:16.8e}'
            for (a, b, c, o, p, w), x in zip(ints.tolist(), vals.tolist())]
    try:
        with open(f'{fname}.dat', 'w') as f:
            f.write(f'{head}\n{Norb} {N_cut} {Nr}\n')
            f.write('\n'.join(body) + '\n')
    except IOError as e:
        print(f'Error: Failed to write {fname}.dat: {e}', flush=True)
        return
    print(f'{label}: Matsubara {Nw_orig} -> {N_cut} | Nr={Nr} | {data_out.nbytes / 1e6:.1f} MB', flush=True)


def _gamma_report(what, data, kmap, invk, full):
    """The pre-transform line of the reference writers: element (0, 0) at the
    lowest frequency, at Gamma and its largest modulus over the grid (every
    grid point is an irreducible point or its time-reversed partner, which
    has the same modulus)."""
    d00 = np.asarray(data)[0, 0, 0]
    p = int(np.nonzero((np.asarray(kmap) == 0).all(axis=1))[0][0])
    if full:
        g = d00[p]
    else:
        g = d00[invk[p, 0] - 1]
        if invk[p, 1] != 0:
            g = np.conj(g)
    print(f'[Pre-FFT] {what}(iw_0): Gamma={g:.6e}, max|..|={np.abs(d00).max():.6e}', flush=True)


def _to_wannier(data, kmap, invk, Nx, Ny, Nz, N_cut, zero_tol, full):
    data_R = flibs.wannier_r(data, kmap, invk, Nx, Ny, Nz, N_cut, full=full)
    return _select_nonzero(np.array(data_R), Nx, Ny, Nz, zero_tol)


def output_self_wannier(sigmak, mu_self, kmap, invk, Nx, Ny, Nz, Nw, temp, N_cut=64, zero_tol=1e-5,
                        fname='self_en_wannier'):
    """Sigma(k, i w_n) [Norb, Norb, Nw, Nk_irr] -> Sigma(R, i w_n) for the first
    N_cut frequencies, written to fname.npz (keys sigma, rvec, iw, mu, temp) and
    fname.dat  (_wannier_io.py:145-176)."""
    _gamma_report('Sigma_11', sigmak, kmap, invk, False)
    data_out, rvec = _to_wannier(sigmak, kmap, invk, Nx, Ny, Nz, N_cut, zero_tol, False)
    if data_out.size == 0:
        print('output_self_wannier: no non-zero R-vectors found', flush=True)
        return
    iw = (2 * np.arange(data_out.shape[2]) + 1) * np.pi * temp
    try:
        np.savez(fname, sigma=data_out, rvec=rvec, iw=iw, mu=mu_self, temp=temp)
    except IOError as e:
        print(f'Error: Failed to write {fname}.npz: {e}', flush=True)
    _write_dat(fname, data_out, rvec, Nw, 'Self-energy', mu=mu_self, temp=temp)


def output_gap_wannier(gap, kmap, invk, Nx, Ny, Nz, Nw, temp, N_cut=64, zero_tol=1e-5, fname='gap_wannier'):
    """Delta(k, i w_n) -> Delta(R, i w_n), written to fname.npz (keys gap, rvec,
    iw, temp) and fname.dat.  gap[Norb, Norb, Nw, Nk] holds the irreducible
    points (unfolded with time reversal) or, when its last axis is Nx Ny Nz,
    every grid point in kmap order  (_wannier_io.py:178-216)."""
    full = np.shape(gap)[-1] == Nx * Ny * Nz
    _gamma_report('Delta_11', gap, kmap, invk, full)
    data_out, rvec = _to_wannier(gap, kmap, invk, Nx, Ny, Nz, N_cut, zero_tol, full)
    if data_out.size == 0:
        print('output_gap_wannier: no non-zero R-vectors found', flush=True)
        return
    iw = (2 * np.arange(data_out.shape[2]) + 1) * np.pi * temp
    try:
        np.savez(fname, gap=data_out, rvec=rvec, iw=iw, temp=temp)
    except IOError as e:
        print(f'Error: Failed to write {fname}.npz: {e}', flush=True)
    _write_dat(fname, data_out, rvec, Nw, 'Gap function', temp=temp)
