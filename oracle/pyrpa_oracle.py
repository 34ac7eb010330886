"""CPU oracle for the FFT-convolved many-body path of pyrpa (TEST INFRASTRUCTURE ONLY).

This module is the checker, never the product: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  ``pyrpa_amd`` (the shipped library and its ctypes host layer)
never imports it and has no CPU fallback.

It restates, in numpy, the algorithms of the reference's Fortran library
``libfmod.so`` on the hot path.  Every public function takes the same
arguments and returns the same array shapes as the reference's Python ctypes
wrapper of the same name (``libs/flibs/*.py``), so a parity test reads like
the reference's own tests.  Internally each function works on the Fortran
index order through ``.T`` views: a Python array of shape (No,No,Nw,Nk) in C
order is the Fortran array ``Gk(Nk,Nw,No,No)``; ``Gk.T[i,j,a,b]`` is the
Fortran element ``Gk(i+1,j+1,a+1,b+1)``.

Third-party arithmetic the reference delegates (not vendored by it, versions
not pinned by it): LAPACK ``zgesv/zgeev/zheev/zgetrf/zgetri/dgesv`` and the
FFTW3 API (``dfftw_plan_dft``: forward sign -1 unnormalised, backward sign +1).
Here they are numpy.linalg (LAPACK) and numpy.fft (pocketfft), which compute
the same mathematical objects.

Pinning (how each function is checked before it is trusted; see tests/):
  * gen_irr_k_TRS            - closed-form Nk of the reference wrapper, k/-k
                               pairing invariants, brute force on tiny meshes
  * gen_ham/get_eig/gen_green0 - closed forms (1-orbital, 2-orbital bands)
  * get_scmat(_orb)          - golden matrices of tests/test_rpa_flex.py:70-113
  * get_chi0_conv/get_chi0   - O(Nk^2) direct sum (algorithm of fchi.f90:920-990),
                               exact Lindhard (tests/_tools.py:245-283)
  * get_chi0_tail            - q=0 sum rule to 1e-10, 2nd-order convergence
                               (tests/test_rpa_flex.py:308-354)
  * get_chis / get_chis_chic / get_vsigma - scalar closed forms, identities
  * calc_sigma               - direct O(Nk^2 Nw^2) convolution sum on tiny grids
  * getinv / gen_green       - G.G^-1 = 1, causality
  * mkself                   - fixed-point residual of the converged Sigma against
                               independently computed (direct-sum) components,
                               n(mu) identity at U=0.  The reference's own tests
                               pin no converged Sigma at U != 0 (parity of the
                               full loop against the reference itself: unpinned)
  * lin_eliash               - lambda=0 at U=0; dense-matrix eigenvalue of the
                               same operator on a tiny grid
"""
from __future__ import annotations

import numpy as np

PI = 3.141592653589793238462643383279  # fmod.f90:4


# --------------------------------------------------------------------------- #
#  k-mesh (finv.f90)
# --------------------------------------------------------------------------- #
def _irr_nk(Nx, Ny, Nz):
    """Closed-form Nk used by the reference wrapper (_hamiltonian.py:367-376)."""
    Nkall = Nx * Ny * Nz
    if Nkall % 2 == 0:
        if Nz % 2 == 0:
            if Nx % 2 == 0 and Ny % 2 == 0:
                return Nkall // 2 + 4
            return Nkall // 2 + 2 if (Nx % 2 == 0 or Ny % 2 == 0) else Nkall // 2 + 1
        return Nkall // 2 + 2 if (Nx % 2 == 0 and Ny % 2 == 0) else Nkall // 2 + 1
    return (Nkall + 1) // 2


def _gen_irr_klist(Nx, Ny, Nz, Nk):
    """Irreducible k-list in the reference's order (finv.f90:102-289).

    Returns integer grid coordinates (Nk,3); klist = coord / (Nx,Ny,Nz)."""
    kl = np.zeros((Nk + 1, 3), dtype=np.int64)  # 1-based slots
    if Nx % 2 == 0 and Ny % 2 == 0:                            # finv.f90:109-133
        for j in range(Ny // 2 + 1):
            for i in range(Nx // 2 + 1):
                kl[i + 1 + (Nx // 2 + 1) * j] = (i, j, 0)
        ini = (Nx // 2 + 1) * (Ny // 2 + 1)
        for j in range(Ny // 2 + 1, Ny):
            for i in range(1, Nx // 2):
                kl[ini + i + (Nx // 2 - 1) * (j - Ny // 2 - 1)] = (i, j, 0)
        ini = Nx * Ny // 2 + 3
    elif (Nx * Ny) % 2 == 0:
        if Nx % 2 != 0:
            # finv.f90:154-155 leaves this branch empty (klist stays zero and
            # iter_k_ini undefined); the reference cannot run it either.
            raise ValueError("odd Nx with even Ny is not supported by the reference k-mesh")
        it = 1
        for j in range((Ny - 1) // 2 + 1):                     # finv.f90:138-152
            for i in range(Nx // 2 + 1):
                kl[it] = (i, j, 0)
                it += 1
        for j in range((Ny + 1) // 2, Ny):
            for i in range(1, Nx // 2):
                kl[it] = (i, j, 0)
                it += 1
        ini = it
    else:                                                      # finv.f90:157-181
        for j in range((Ny - 1) // 2 + 1):
            for i in range((Nx - 1) // 2 + 1):
                kl[i + 1 + ((Nx + 1) // 2) * j] = (i, j, 0)
        ini = (Nx + 1) * (Ny + 1) // 4
        for j in range((Ny + 1) // 2, Ny):
            for i in range(1, (Nx - 1) // 2 + 1):
                kl[ini + i + ((Nx - 1) // 2) * (j - (Ny + 1) // 2)] = (i, j, 0)
        ini = (Nx * Ny + 1) // 2 + 1
    if Nz > 1:
        if Nz % 2 == 0:                                        # finv.f90:186-271
            for k in range(1, Nz // 2):
                for j in range(Ny):
                    for i in range(Nx):
                        kl[ini + i + Nx * j + Nx * Ny * (k - 1)] = (i, j, k)
            ini = ini + Nx * Ny * (Nz // 2 - 1)
            h = Nz // 2  # kz = 0.5
            if Nx % 2 == 0 and Ny % 2 == 0:
                for j in range(Ny // 2 + 1):
                    for i in range(Nx // 2 + 1):
                        kl[ini + i + (Nx // 2 + 1) * j] = (i, j, h)
                ini = ini + (Nx // 2 + 1) * (Ny // 2 + 1) - 1
                for j in range(Ny // 2 + 1, Ny):
                    for i in range(1, Nx // 2):
                        kl[ini + i + (Nx // 2 - 1) * (j - Ny // 2 - 1)] = (i, j, h)
            else:
                it = ini
                if (Nx * Ny) % 2 == 0:
                    for j in range((Ny - 1) // 2 + 1):
                        for i in range(Nx // 2 + 1):
                            kl[it] = (i, j, h)
                            it += 1
                    for j in range((Ny + 1) // 2, Ny):
                        for i in range(1, Nx // 2):
                            kl[it] = (i, j, h)
                            it += 1
                else:
                    for j in range((Ny - 1) // 2 + 1):
                        for i in range((Nx - 1) // 2 + 1):
                            kl[it] = (i, j, h)
                            it += 1
                    for j in range((Ny + 1) // 2, Ny):
                        for i in range(1, (Nx - 1) // 2 + 1):
                            kl[it] = (i, j, h)
                            it += 1
        else:                                                  # finv.f90:272-286
            for k in range(1, (Nz - 1) // 2 + 1):
                for j in range(Ny):
                    for i in range(Nx):
                        kl[ini + i + Nx * j + Nx * Ny * (k - 1)] = (i, j, k)
    return kl[1:]


def gen_irr_k_TRS(Nx, Ny, Nz):
    """klist[Nk,3], kmap[Nkall,3], invk[Nkall,3]  (finv.f90:1-290,
    wrapper _hamiltonian.py:356-382).

    invk[i] = (irreducible index (1-based), 0: k==k_irr / 1: k==-k_irr mod 1,
    1-based index of -k in the full list).  The full-BZ search follows the
    reference loop order: for each full k, the first irreducible j (in list
    order) whose k or -k matches, direct match tested first."""
    Nk = _irr_nk(Nx, Ny, Nz)
    Nkall = Nx * Ny * Nz
    kc = _gen_irr_klist(Nx, Ny, Nz, Nk)
    N = np.array([Nx, Ny, Nz])
    klist = kc / N
    ii = np.arange(Nkall)
    kmap = np.stack([ii % Nx, (ii // Nx) % Ny, ii // (Nx * Ny)], axis=1).astype(np.int64)
    neg = (-kc) % N
    invk = np.zeros((Nkall, 3), dtype=np.int64)
    for i in range(Nkall):
        d = np.all(kc == kmap[i], axis=1)
        n = np.all(neg == kmap[i], axis=1)
        hit = np.nonzero(d | n)[0]
        if len(hit):
            j = hit[0]
            invk[i, 0] = j + 1
            invk[i, 1] = 0 if d[j] else 1
    mk = (-kmap) % N
    invk[:, 2] = mk[:, 0] + Nx * mk[:, 1] + Nx * Ny * mk[:, 2] + 1
    return klist, kmap, invk


def kweight(invk, Nk):
    """Full-BZ multiplicity of each irreducible k (finv.f90:292-305)."""
    w = np.zeros(Nk)
    np.add.at(w, invk[:, 0] - 1, 1.0)
    return w


# --------------------------------------------------------------------------- #
#  Hamiltonian / bands / Green functions (fmod.f90, fself.f90)
# --------------------------------------------------------------------------- #
def gen_ham(klist, ham_r, rvec):
    """H(k) = sum_R H(R) exp(-2 pi i k.R), lower triangle then Hermitised
    (fmod.f90:162-207).  Python shapes: ham_r[Nr,No,No], hamk[Nk,No,No]."""
    Hr = np.asarray(ham_r).T                      # Fortran ham_r(m,l,R)
    ph = 2 * PI * (np.asarray(klist) @ np.asarray(rvec).T)   # [Nk,Nr]
    cph = np.cos(ph) - 1j * np.sin(ph)
    Hk = np.einsum('mlr,kr->mlk', Hr, cph)        # Fortran ham_k(m,l,k)
    No = Hk.shape[0]
    for l in range(No):
        Hk[l, l] = Hk[l, l].real
        for m in range(l + 1, No):
            Hk[l, m] = np.conj(Hk[m, l])
    return np.ascontiguousarray(Hk.T)


def get_eig(hamk):
    """eig[Nk,No], uni[Nk,No,No]; Fortran uni(:,n,k) is eigenvector n
    (fmod.f90:209-248, zheev 'V','U')."""
    Hf = np.asarray(hamk).T                        # (No,No,Nk): Hf[m,l,k] = ham_k(m,l,k)
    A = np.moveaxis(Hf, 2, 0)                      # A[k] = matrix rows m, cols l
    w, v = np.linalg.eigh(A, UPLO='U')
    return w, np.ascontiguousarray(np.transpose(v, (0, 2, 1)))


def get_ffermi(eig, mu, temp):
    """f = 1/2 - 1/2 tanh((e-mu)/2T) (fmod.f90:312-340)."""
    t = max(temp, 1e-12)
    return 0.5 - 0.5 * np.tanh((np.asarray(eig) - mu) * (0.5 / t))


def calc_mu(eig, Nk, fill, temp):
    """mu with sum_k,n f(e - mu) = fill Nk by bisection between the band edges,
    clamped to the closer edge when the filling lies outside the band range
    (libs/plibs/_bands.py:50-78, which uses brentq on the same function)."""
    e = np.asarray(eig, dtype=np.float64).ravel()
    func = lambda mu: fill * Nk - get_ffermi(e, mu, temp).sum()
    lo, hi = float(e.min()), float(e.max())
    flo, fhi = func(lo), func(hi)
    if flo * fhi > 0:
        return lo if abs(flo) < abs(fhi) else hi
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        fm = func(mid)
        if (fm > 0) == (flo > 0):
            lo, flo = mid, fm
        else:
            hi = mid
        if hi - lo < 1e-15 * max(1.0, abs(mid)):
            break
    return 0.5 * (lo + hi)


def get_mu(ham_r, rvec, temp, fill, kmesh=40):
    """Initial mu on a kmesh^3 full mesh (libs/plibs/_calc.py:298-309 with
    get_emesh, _bands.py:325-344)."""
    g = np.arange(kmesh) / kmesh
    kx, ky, kz = np.meshgrid(g, g, g, indexing='ij')
    kl = np.stack([kx.ravel(), ky.ravel(), kz.ravel()], axis=1)
    eig, _ = get_eig(gen_ham(kl, ham_r, rvec))
    return calc_mu(eig, len(kl), fill, temp)


def gen_Green0(eig, uni, mu, temp, Nw):
    """G0_{ml}(k,iw_j) = sum_n U_mn U*_ln / (i w_j + mu - e_n), w_j=(2j-1) pi T
    (fself.f90:1-41).  Returns Python Gk[No,No,Nw,Nk]."""
    ef = np.asarray(eig).T                          # eig(n,k)
    Uf = np.asarray(uni).T                          # uni(m,n,k)
    wl = (2 * np.arange(1, Nw + 1) - 1) * PI * temp
    iw = mu + 1j * wl                                # [Nw]
    den = 1.0 / (iw[None, None, :] - ef[:, :, None])  # [n,k,w]
    G = np.einsum('mnk,lnk,nkw->kwml', Uf, np.conj(Uf), den)
    return np.ascontiguousarray(G.T)


def gen_green_inv(selfen, hamk, mu, temp):
    """G^-1 = (iw+mu) - H(k) - Sigma (fself.f90:43-83); Python layouts."""
    Sf = np.asarray(selfen).T                       # (Nk,Nw,No,No)
    Hf = np.asarray(hamk).T                         # (No,No,Nk)
    Nk, Nw, No, _ = Sf.shape
    wl = (2 * np.arange(1, Nw + 1) - 1) * PI * temp
    Gi = -(np.transpose(Hf, (2, 0, 1))[:, None, :, :] + Sf)
    for l in range(No):
        Gi[:, :, l, l] += (mu + 1j * wl)[None, :]
    return np.ascontiguousarray(Gi.T)


def getinv(Gk):
    """Per-(k,iw) matrix inverse (zgetrf+zgetri, fself.f90:131-157)."""
    Gf = np.asarray(Gk).T
    return np.ascontiguousarray(np.linalg.inv(Gf).T)


def gen_green(selfen, hamk, mu, temp):
    """G = [iw + mu - H - Sigma]^-1 (_green.py:99-117)."""
    return getinv(gen_green_inv(selfen, hamk, mu, temp))


# --------------------------------------------------------------------------- #
#  orbital-pair bookkeeping and vertices (fUJ.f90, _response.py, fchi.f90)
# --------------------------------------------------------------------------- #
def get_chi_orb_list(Norb, site_prof):
    """olist (Nchi,2) with olist[:,0] running fastest (_response.py:382-413).
    Single site: returned as the same F-contiguous transposed view the
    reference produces, so its memory is Fortran olist(Nchi,2)."""
    if len(site_prof) == 1:
        tmp = np.arange(Norb) + 1
        o1, o2 = np.meshgrid(tmp, tmp)
        ol = np.array([o1.flatten(), o2.flatten()]).T
        site = np.ones(len(ol), dtype=np.int64)
        return ol.astype(np.int64), site
    raise NotImplementedError("multi-site olist is C-contiguous in the reference (misread by Fortran)")


def olist_f(olist):
    """Fortran view olist(Nchi,2) as an (Nchi,2) int array (0-based orbitals).

    ctypes passes the data pointer of the numpy array; Fortran reads it as
    column-major (Nchi,2), i.e. memory = [olist(:,1), olist(:,2)]."""
    ol = np.asarray(olist, dtype=np.int64)
    raw = ol.ravel(order='K')          # elements in memory order, as ctypes hands them over
    return raw.reshape((2, ol.shape[0])).T - 1


def get_scmat(olist, site, U, J):
    """Kanamori S, C (fUJ.f90:1-57). Python shapes (Nchi,Nchi); Fortran
    Smat(j,i) is Python Smat[i,j]."""
    ol = olist_f(olist) + 1
    Nchi = len(ol)
    S = np.zeros((Nchi, Nchi))
    C = np.zeros((Nchi, Nchi))
    Up = U - 2 * J
    for i in range(Nchi):
        for j in range(Nchi):
            if site[i] != site[j]:
                continue
            if ol[i, 0] == ol[i, 1] and ol[j, 0] == ol[j, 1]:
                if ol[i, 0] == ol[j, 0]:
                    S[i, j], C[i, j] = U, U
                else:
                    S[i, j], C[i, j] = J, 2 * Up - J
            elif ol[i, 0] == ol[j, 0] and ol[i, 1] == ol[j, 1]:
                S[i, j], C[i, j] = Up, -Up + 2 * J
            elif ol[i, 0] == ol[j, 1] and ol[i, 1] == ol[j, 0]:
                S[i, j], C[i, j] = J, J
    return S, C


def get_scmat_orb(olist, site, Umat, Jmat):
    """Orbital-dependent S, C (fUJ.f90:59-107)."""
    ol = olist_f(olist)
    Uf = np.asarray(Umat).T
    Jf = np.asarray(Jmat).T
    Nchi = len(ol)
    S = np.zeros((Nchi, Nchi))
    C = np.zeros((Nchi, Nchi))
    for i in range(Nchi):
        for j in range(Nchi):
            if site[i] != site[j]:
                continue
            a, b = ol[i]
            c, d = ol[j]
            if a == b and c == d:
                if a == c:
                    S[i, j] = C[i, j] = Uf[a, a]
                else:
                    S[i, j] = Jf[a, c]
                    C[i, j] = 2 * Uf[a, c] - Jf[a, c]
            elif a == c and b == d:
                S[i, j] = Uf[a, b]
                C[i, j] = -Uf[a, b] + 2 * Jf[a, b]
            elif a == d and b == c:
                S[i, j] = C[i, j] = Jf[a, b]
    return S, C


def get_chi_map(olist):
    """chi_map(Nchi,Nchi,2), irr_chi(P,2) of fchi.f90:421-481, returned as
    0-based Fortran-index numpy arrays cmap[m1,l1,c], irr[i,c]."""
    ol = olist_f(olist)
    Nchi = len(ol)
    cmap = np.zeros((Nchi, Nchi, 2), dtype=np.int64)
    cirr = np.zeros((Nchi, Nchi), dtype=np.int64)
    for l1 in range(Nchi):
        for m1 in range(Nchi):
            t1 = (ol[l1, 0], ol[l1, 1], ol[m1, 0], ol[m1, 1])
            found = False
            for l2 in range(Nchi):
                for m2 in range(Nchi):
                    if t1 == (ol[m2, 1], ol[m2, 0], ol[l2, 1], ol[l2, 0]):
                        cmap[m1, l1] = (m2, l2)
                        if cirr[m2, l2] == 0:
                            cirr[m1, l1] = 1
                        found = True
                        break
                if found:
                    break
    irr = [(l1, m1) for l1 in range(Nchi) for m1 in range(Nchi) if cirr[m1, l1] == 1]
    return cmap, np.array(irr, dtype=np.int64).reshape(-1, 2)


# --------------------------------------------------------------------------- #
#  full-grid scatter / FFT conventions (fmod.f90:67-160, fchi.f90:563-583)
# --------------------------------------------------------------------------- #
# host threads of the oracle's transforms (scipy.fft workers; pocketfft, the
# library the reference vendors); 1 unless a timing harness raises it
FFT_WORKERS = 1


def _fftn(a, axes):
    import scipy.fft
    return scipy.fft.fftn(a, axes=axes, workers=FFT_WORKERS)


def fft4(a):
    """FFT(.true.): 4D, sign -1, unnormalised (fmod.f90:67-116)."""
    return _fftn(a, (0, 1, 2, 3))


def ifft4(a):
    """FFT(.false.): 4D, sign +1, divided by the grid size."""
    import scipy.fft
    return scipy.fft.ifftn(a, axes=(0, 1, 2, 3), workers=FFT_WORKERS)


def _spatial(kmap, Nx, Ny, Nz):
    """Flat index of each full-BZ point in a C-order (Nx,Ny,Nz) block."""
    return (kmap[:, 0] * Ny + kmap[:, 1]) * Nz + kmap[:, 2]


def _full_green(Gf, a, b, kmap, invk, Nx, Ny, Nz):
    """Full-BZ, full-frequency G_ab(k, slot) from irreducible storage with
    TRS rules (fchi.f90:563-583): slot j -> w_j, slot M-1-j -> -w_j;
    k = -k_irr uses G(-k) = G(k)^T; -w uses G(k,-w) = G(k,w)^dagger."""
    Nk, Nw = Gf.shape[:2]
    M = 2 * Nw
    out = np.zeros((Nx * Ny * Nz, M), dtype=complex)
    irr = invk[:, 0] - 1
    fl = invk[:, 1]
    sp = _spatial(kmap, Nx, Ny, Nz)
    d = fl == 0
    r = fl == 1
    pos = Gf[:, :, a, b]
    tr = Gf[:, :, b, a]
    out[sp[d], :Nw] = pos[irr[d]]
    out[sp[d], Nw:] = np.conj(tr[irr[d]])[:, ::-1]
    out[sp[r], :Nw] = tr[irr[r]]
    out[sp[r], Nw:] = np.conj(pos[irr[r]])[:, ::-1]
    return out.reshape(Nx, Ny, Nz, M)


def _full_boson(Vf, m, l, static, kmap, invk, Nx, Ny, Nz):
    """Full-grid bosonic V_(m,l)(q, slot) (fself.f90:203-224): slot 0 = nu_0,
    slots j, M-j = +-nu_j, slot Nw = the nu -> infinity value `static`."""
    Nk, Nw = Vf.shape[:2]
    M = 2 * Nw
    out = np.zeros((Nx * Ny * Nz, M), dtype=complex)
    irr = invk[:, 0] - 1
    fl = invk[:, 1]
    sp = _spatial(kmap, Nx, Ny, Nz)
    d = fl == 0
    r = fl == 1
    vml = Vf[:, :, m, l]
    vlm = Vf[:, :, l, m]
    out[sp[d], :Nw] = vml[irr[d]]
    out[sp[r], :Nw] = vlm[irr[r]]
    out[:, Nw] = static
    out[sp[d], Nw + 1:] = np.conj(vlm[irr[d]][:, 1:])[:, ::-1]
    out[sp[r], Nw + 1:] = np.conj(vml[irr[r]][:, 1:])[:, ::-1]
    return out.reshape(Nx, Ny, Nz, M)


def _gather(full, kmap, invk, Nk, Nw):
    """Irreducible (Nk, Nw) slice of a full grid (flag-0 points, first Nw slots)."""
    Nx, Ny, Nz, M = full.shape
    d = invk[:, 1] == 0
    out = np.zeros((Nk, Nw), dtype=complex)
    out[invk[d, 0] - 1] = full.reshape(-1, M)[_spatial(kmap, Nx, Ny, Nz)[d], :Nw]
    return out


def _reverse(a):
    """a(-r,-tau) on the periodic grid (fchi.f90:532-556)."""
    idx = [(-np.arange(n)) % n for n in a.shape]
    return a[np.ix_(*idx)]


# --------------------------------------------------------------------------- #
#  chi0 (fchi.f90)
# --------------------------------------------------------------------------- #
def _chi0_conv_acc(chif, Gf, kmap, invk, olist, temp, coef, Nx, Ny, Nz):
    """chi += coef * FFT bubble (fchi.f90:508-613), in place on Fortran view."""
    ol = olist_f(olist)
    cmap, irr = get_chi_map(olist)
    Nk, Nw = Gf.shape[:2]
    Nkall = len(kmap)
    weight = coef * temp / Nkall
    for l, m in irr:
        g13 = fft4(_full_green(Gf, ol[l, 0], ol[m, 0], kmap, invk, Nx, Ny, Nz))
        g42 = fft4(_full_green(Gf, ol[m, 1], ol[l, 1], kmap, invk, Nx, Ny, Nz))
        tmp = ifft4(-g13 * _reverse(g42))
        chif[:, :, m, l] += _gather(tmp, kmap, invk, Nk, Nw) * weight
        pm, pl = cmap[m, l]
        chif[:, :, pm, pl] = np.conj(chif[:, :, m, l])


def _threshold(a, eps=1e-9):
    """Zero |Re| or |Im| below eps, Re first (fchi.f90:483-506, fself.f90:272-284)."""
    re = np.where(np.abs(a.real) < eps, 0.0, a.real)
    im = np.where(np.abs(a.imag) < eps, 0.0, a.imag)
    return re + 1j * im


def get_chi0_conv(Gk, kmap, invk, olist, temp, Nx, Ny, Nz):
    """chi0 by sharp-cutoff FFT convolution (fchi.f90:615-635,
    wrapper _susceptibility.py:244-269).  Returns chi[Nchi,Nchi,Nw,Nk]."""
    Gf = np.asarray(Gk).T
    Nk, Nw = Gf.shape[:2]
    Nchi = len(olist)
    chif = np.zeros((Nk, Nw, Nchi, Nchi), dtype=complex)
    _chi0_conv_acc(chif, Gf, kmap, invk, olist, temp, 1.0, Nx, Ny, Nz)
    chif = _threshold(chif)
    return np.ascontiguousarray(chif.T)


def ckchi(chi, Smat, Cmat):
    """Stoner factor max_q max Re eig(chi0(q,0) S) (fchi.f90:352-399)."""
    chif = np.asarray(chi).T
    S = np.asarray(Smat).T
    C = np.asarray(Cmat).T
    best = -1e5
    for i in range(chif.shape[0]):
        A = chif[i, 0]
        v = np.linalg.eigvals(A @ S).real.max()
        if v > best:
            best = v
    return best


def get_chi0(Smat, Cmat, Gk, olist, kmap, invk, temp, Nx, Ny, Nz):
    """chi0 + Stoner factor (fchi.f90:401-419, _susceptibility.py:432-458)."""
    chi = get_chi0_conv(Gk, kmap, invk, olist, temp, Nx, Ny, Nz)
    return chi, ckchi(chi, Smat, Cmat)


def _chi0_ref_tau(chif, eig, uni, mu, kmap, invk, olist, temp, Nx, Ny, Nz, Nw):
    """Analytic reference bubble chi0[G0,G0] on the tau grid (fchi.f90:637-850)."""
    ol = olist_f(olist)
    cmap, irr = get_chi_map(olist)
    ef = np.asarray(eig).T                 # (No,Nk)
    Uf = np.asarray(uni).T                 # (No,No,Nk)
    No, Nk = ef.shape
    Nkall = len(kmap)
    M = 2 * Nw
    beta = 1.0 / temp
    dtau = beta / M
    weight = 1.0 / (temp * Nkall)
    xi = ef - mu
    tau = np.arange(M) * dtau
    pos = xi >= 0
    fn = np.where(pos, np.exp(-beta * np.abs(xi)) / (1 + np.exp(-beta * np.abs(xi))),
                  1.0 / (1.0 + np.exp(-beta * np.abs(xi))))
    fn = np.where(pos, fn, 1.0 / (1.0 + np.exp(beta * xi)))
    X = xi[None]
    T_ = tau[:, None, None]
    gp = np.where(pos[None], -np.exp(-X * T_) / (1 + np.exp(-beta * X)),
                  -np.exp(X * (beta - T_)) / (1 + np.exp(beta * X)))
    gm = np.where(pos[None], np.exp(-X * (beta - T_)) / (1 + np.exp(-beta * X)),
                  np.exp(X * T_) / (1 + np.exp(beta * X)))
    # rho(m,l,k) = sum_n U_mn U*_ln f_n ; rhobar over full BZ (transpose for -k)
    rho = np.einsum('mnk,lnk,nk->mlk', Uf, np.conj(Uf), fn)
    rb = np.zeros((No, No), dtype=complex)
    for i in range(Nkall):
        kk = invk[i, 0] - 1
        if invk[i, 1] == 0:
            rb += rho[:, :, kk]
        elif invk[i, 1] == 1:
            rb += rho[:, :, kk].T
    rb /= Nkall
    sawc = np.zeros(Nw, dtype=complex)
    s = np.arange(M)
    for j in range(1, Nw + 1):
        if j > 1:
            nu = 2.0 * PI * (j - 1) * temp
            sawc[j - 1] = 1j / nu
        ph = np.exp(1j * 2.0 * PI * (j - 1) * s / M)
        sawc[j - 1] -= np.sum(dtau * ph * (0.5 - s / M))
    irrk = invk[:, 0] - 1
    fl = invk[:, 1]
    for l, m in irr:
        a13, b13 = ol[l, 0], ol[m, 0]
        a42, b42 = ol[m, 1], ol[l, 1]
        U13 = Uf[a13] * np.conj(Uf[b13])            # (n,k)
        U42 = Uf[a42] * np.conj(Uf[b42])
        c13 = np.where(fl[None, :] == 0, U13[:, irrk], np.conj(U13[:, irrk]))  # (n,Nkall)
        c42 = np.where(fl[None, :] == 0, U42[:, irrk], np.conj(U42[:, irrk]))
        zA = np.einsum('nf,snf->fs', c13, gp[:, :, irrk])
        zB = np.einsum('nf,snf->fs', c42, gm[:, :, irrk])
        sp = _spatial(kmap, Nx, Ny, Nz)
        A = np.zeros((Nx * Ny * Nz, M), dtype=complex)
        B = np.zeros((Nx * Ny * Nz, M), dtype=complex)
        A[sp] = zA
        B[sp] = zB
        A = A.reshape(Nx, Ny, Nz, M)
        B = B.reshape(Nx, Ny, Nz, M)
        A = _fftn(A, (0, 1, 2))
        B = _fftn(B, (0, 1, 2))
        idx = [(-np.arange(n)) % n for n in (Nx, Ny, Nz)]
        Brev = B[np.ix_(idx[0], idx[1], idx[2], np.arange(M))]
        tmp = ifft4(-A * Brev)
        J = 0j
        if a13 == b13:
            J += rb[a42, b42]
        if a42 == b42:
            J -= rb[a13, b13]
        chif[:, :, m, l] += _gather(tmp, kmap, invk, Nk, Nw) * weight + J * sawc[None, :]
        pm, pl = cmap[m, l]
        chif[:, :, pm, pl] = np.conj(chif[:, :, m, l])


def get_chi0_tail(Smat, Cmat, Gk, eig, uni, olist, kmap, invk, mu, temp, Nx, Ny, Nz):
    """Tail-corrected chi0 = conv[G] - conv[G0] + ref[G0] (fchi.f90:852-918,
    _susceptibility.py:460-494)."""
    Gf = np.asarray(Gk).T
    Nk, Nw = Gf.shape[:2]
    Nchi = len(olist)
    chif = np.zeros((Nk, Nw, Nchi, Nchi), dtype=complex)
    _chi0_conv_acc(chif, Gf, kmap, invk, olist, temp, 1.0, Nx, Ny, Nz)
    G0f = np.asarray(gen_Green0(eig, uni, mu, temp, Nw)).T
    _chi0_conv_acc(chif, G0f, kmap, invk, olist, temp, -1.0, Nx, Ny, Nz)
    _chi0_ref_tau(chif, eig, uni, mu, kmap, invk, olist, temp, Nx, Ny, Nz, Nw)
    chif = _threshold(chif)
    chi = np.ascontiguousarray(chif.T)
    return chi, ckchi(chi, Smat, Cmat)


def get_chi0_sum(Gk, klist, kmap, invk, olist, temp, Nx, Ny, Nz):
    """O(Nk^2) direct Matsubara/BZ sum chi_{(l1l2)(m1m2)}(q,nu_m) =
    -(T/Nkall) sum_{k,n} G13(k+q, w_n+nu_m) G42(k, w_n) over the same truncated
    2Nw-frequency circular grid (the algorithm of fchi.f90:920-990).  An
    independent check of the FFT convolution, not a restatement of it."""
    Gf = np.asarray(Gk).T
    Nk, Nw = Gf.shape[:2]
    ol = olist_f(olist)
    cmap, irr = get_chi_map(olist)
    Nkall = len(kmap)
    M = 2 * Nw
    chif = np.zeros((Nk, Nw, len(olist), len(olist)), dtype=complex)
    d = invk[:, 1] == 0
    for l, m in irr:
        g13 = _full_green(Gf, ol[l, 0], ol[m, 0], kmap, invk, Nx, Ny, Nz)
        g42 = _full_green(Gf, ol[m, 1], ol[l, 1], kmap, invk, Nx, Ny, Nz)
        for i in np.nonzero(d)[0]:
            q = kmap[i]
            sh = np.roll(g13, shift=(-q[0], -q[1], -q[2]), axis=(0, 1, 2))   # G13(k+q)
            for mm in range(Nw):
                g = np.roll(sh, shift=-mm, axis=3)                           # w_n + nu_m
                chif[invk[i, 0] - 1, mm, m, l] = -(g * g42).sum() * temp / Nkall
        pm, pl = cmap[m, l]
        chif[:, :, pm, pl] = np.conj(chif[:, :, m, l])
    return np.ascontiguousarray(chif.T)


# --------------------------------------------------------------------------- #
#  RPA algebra (fchi.f90:319-350, fUJ.f90:253-365, fself.f90:783-850)
# --------------------------------------------------------------------------- #
def get_chis(chi0, Smat):
    """chi_s(nu) = (I - chi0 S)^-1 chi0 at one q (fchi.f90:319-350)."""
    c = np.asarray(chi0).T                         # (Nchi,Nchi,Nw)
    S = np.asarray(Smat).T
    n = c.shape[0]
    out = np.zeros_like(c)
    for i in range(c.shape[2]):
        A = c[:, :, i]
        out[:, :, i] = np.linalg.solve(np.eye(n) - A @ S, A)
    return np.ascontiguousarray(out.T)


def get_chis_chic(chi, Smat, Cmat):
    """Static RPA chi_s, chi_c at all irreducible q (fself.f90:783-816)."""
    chif = np.asarray(chi).T
    S = np.asarray(Smat).T
    C = np.asarray(Cmat).T
    Nk = chif.shape[0]
    n = S.shape[0]
    A = chif[:, 0]
    chis = np.linalg.solve(np.eye(n)[None] - A @ S, A)
    chic = np.linalg.solve(np.eye(n)[None] + A @ C, A)
    return np.ascontiguousarray(chis.T), np.ascontiguousarray(chic.T)


def get_chis_chic_all(chi, Smat, Cmat):
    """chi_s = (I - chi0 S)^-1 chi0 and chi_c = (I + chi0 C)^-1 chi0 at every
    (q, i nu): the algebra of get_chis (fchi.f90:319-350) and get_chis_chic
    (fself.f90:783-816) over the whole batch, on the Fortran matrices
    chi(k,w,:,:).  chi: Python [Nchi,Nchi,Nw,Nk]; returns the same shape twice."""
    A = np.asarray(chi).T                      # (Nk,Nw,Nchi,Nchi) = Fortran chi(k,w,:,:)
    S = np.asarray(Smat).T
    C = np.asarray(Cmat).T
    eye = np.eye(S.shape[0])[None, None]
    chis = np.linalg.solve(eye - A @ S, A)
    chic = np.linalg.solve(eye + A @ C, A)
    return np.ascontiguousarray(chis.T), np.ascontiguousarray(chic.T)


def get_eig_or_tr_chi(chi, invk, sw_eig):
    """Leading Re eigenvalue or trace per full-BZ k (fself.f90:818-850)."""
    chif = np.asarray(chi).T                       # (Nk,Nchi,Nchi)
    sel = chif[invk[:, 0] - 1]
    if sw_eig:
        return np.linalg.eigvals(sel).real.max(axis=1).astype(complex)
    return np.trace(sel, axis1=1, axis2=2)


def _vsigma_f(chif, S, C):
    A = chif
    n = S.shape[0]
    I = np.eye(n)
    AS = A @ S
    AC = A @ C
    Xs = np.linalg.solve(I - AS, AS)
    Xc = np.linalg.solve(I + AC, AC)
    return (1.5 * S - 0.5 * C) + 1.5 * (S @ Xs) + 0.5 * (C @ Xc) - 0.25 * ((S + C) @ (AS + AC))


def get_Vsigma_nosoc_flex(chi, Smat, Cmat):
    """FLEX vertex V_sigma (fUJ.f90:318-365, _selfenergy.py:62-74)."""
    chif = np.asarray(chi).T
    V = _vsigma_f(chif, np.asarray(Smat).T, np.asarray(Cmat).T)
    return np.ascontiguousarray(V.T)


def get_V_delta(chi, Smat, Cmat, singlet=True):
    """Pairing vertex V_Delta (fUJ.f90:253-316)."""
    chif = np.asarray(chi).T
    S = np.asarray(Smat).T
    C = np.asarray(Cmat).T
    n = S.shape[0]
    I = np.eye(n)
    AS = chif @ S
    AC = chif @ C
    Xs = np.linalg.solve(I - AS, AS)
    Xc = np.linalg.solve(I + AC, AC)
    if singlet:
        V = 0.5 * (S + C) + 1.5 * (S @ Xs) - 0.5 * (C @ Xc)
    else:
        V = 0.5 * (S - C) - 0.5 * (S @ Xs) - 0.5 * (C @ Xc)
    return np.ascontiguousarray(V.T)


# --------------------------------------------------------------------------- #
#  self-energy (fself.f90:159-285)
# --------------------------------------------------------------------------- #
def calc_sigma(Gk, Vsigma, Smat, Cmat, kmap, invk, olist, temp, Nx, Ny, Nz):
    """Sigma = (T/Nkall) V (*) G over all Nchi^2 pairs, then 1e-9 threshold."""
    Gf = np.asarray(Gk).T
    Vf = np.asarray(Vsigma).T
    S = np.asarray(Smat).T
    C = np.asarray(Cmat).T
    ol = olist_f(olist)
    Nk, Nw, No, _ = Gf.shape
    Nchi = len(ol)
    Nkall = len(kmap)
    sig = np.zeros((Nk, Nw, No, No), dtype=complex)
    gft = {}
    for l in range(Nchi):
        for m in range(Nchi):
            v = fft4(_full_boson(Vf, m, l, 1.5 * S[m, l] - 0.5 * C[m, l], kmap, invk, Nx, Ny, Nz))
            key = (ol[m, 1], ol[l, 1])
            if key not in gft:
                gft[key] = fft4(_full_green(Gf, key[0], key[1], kmap, invk, Nx, Ny, Nz))
            tmp = ifft4(v * gft[key])
            sig[:, :, ol[m, 0], ol[l, 0]] += _gather(tmp, kmap, invk, Nk, Nw)
    sig *= temp / Nkall
    return np.ascontiguousarray(_threshold(sig).T)


def calc_sigma_direct(Gk, Vsigma, Smat, Cmat, kmap, invk, olist, temp, Nx, Ny, Nz):
    """Direct circular-convolution sum of the same Sigma (independent check):
    Sigma_{m1 l1}(k, w_n) = (T/Nkall) sum_{q,m} V(q, nu_m) G(k-q, w_n - nu_m)."""
    Gf = np.asarray(Gk).T
    Vf = np.asarray(Vsigma).T
    S = np.asarray(Smat).T
    C = np.asarray(Cmat).T
    ol = olist_f(olist)
    Nk, Nw, No, _ = Gf.shape
    Nkall = len(kmap)
    M = 2 * Nw
    sig = np.zeros((Nk, Nw, No, No), dtype=complex)
    d = np.nonzero(invk[:, 1] == 0)[0]
    for l in range(len(ol)):
        for m in range(len(ol)):
            v = _full_boson(Vf, m, l, 1.5 * S[m, l] - 0.5 * C[m, l], kmap, invk, Nx, Ny, Nz)
            g = _full_green(Gf, ol[m, 1], ol[l, 1], kmap, invk, Nx, Ny, Nz)
            for i in d:
                kx, ky, kz = kmap[i]
                gs = g[(kx - np.arange(Nx)) % Nx][:, (ky - np.arange(Ny)) % Ny][:, :, (kz - np.arange(Nz)) % Nz]
                for n in range(Nw):
                    gw = gs[..., (n - np.arange(M)) % M]
                    sig[invk[i, 0] - 1, n, ol[m, 0], ol[l, 0]] += (v * gw).sum()
    sig *= temp / Nkall
    return np.ascontiguousarray(sig.T)


# --------------------------------------------------------------------------- #
#  FLEX loop (fself.f90:287-740)
# --------------------------------------------------------------------------- #
class _Flex:
    """State of one mkself call; mirrors the contained subroutines of mkself."""

    def __init__(self, hamk, eig, rfill, temp, m_diis, No, Nk, Nw):
        self.Hf = np.moveaxis(np.asarray(hamk).T, 2, 0)      # [k][p,q] = ham_k(p,q,k)
        self.ef = np.asarray(eig).T                           # eig(n,k)
        self.rfill = rfill
        self.temp = temp
        self.m_diis = m_diis
        self.Nk, self.Nw, self.No = Nk, Nw, No
        self.n_hist = 0
        self.i_hist = 0
        self.xout = [None] * m_diis
        self.res = [None] * m_diis
        self.bracket_valid = False
        self.muS_cache = self.muL_cache = 0.0
        self.esterr = 0.0
        wl = (2 * np.arange(1, Nw + 1) - 1) * PI * temp
        self.iw = 1j * wl

    def precompute(self, sigf):
        A = self.Hf[:, None] + sigf                           # [k,w,p,q]
        self.lam = np.linalg.eigvals(A)                        # [k,w,n]

    def get_rn(self, rmu):
        """n(mu) (fself.f90:687-714): unweighted over the irreducible set."""
        t = np.sum(0.5 * (1.0 - np.tanh(0.5 * (self.ef - rmu) / self.temp)))
        z = rmu + self.iw[None, :, None]
        dg = np.sum((1.0 / (z - self.lam) - 1.0 / (z - self.ef.T[:, None, :])).real)
        return (t + 2 * self.temp * dg) / self.Nk

    def renew_mu(self, mu, mu_old, sigf):
        """Bracket + Brent/IQI + final secant (fself.f90:544-685)."""
        self.precompute(sigf)
        eps = 1.0e-8 if self.esterr > 1.0e-2 else self.esterr * 0.1
        dmu = abs(mu - mu_old) * 2.0
        if dmu < eps * 4.0:
            dmu = eps * 4.0
        rfill = self.rfill
        found = False
        if self.bracket_valid:
            rnS = self.get_rn(self.muS_cache)
            rnL = self.get_rn(self.muL_cache)
            if rnS < rfill and rnL > rfill:
                muS, muL = self.muS_cache, self.muL_cache
                found = True
        if not found:
            muL = mu + dmu
            muS = mu - dmu
            for it in range(1, 101):
                mu = muL
                rnL = self.get_rn(mu)
                if rnL < rfill:
                    if abs(rfill - rnL) > 0.5:
                        muL += 1.0
                    elif abs(rfill - rnL) > dmu:
                        muL += 0.5
                    else:
                        muL += dmu
                else:
                    break
                if it == 100:
                    raise RuntimeError("Too many")
            for it in range(1, 101):
                mu = muS
                rnS = self.get_rn(mu)
                if rnS > rfill:
                    if abs(rnS - rfill) > 0.5:
                        muS -= 1.0
                    elif abs(rnS - rfill) > dmu:
                        muS -= 0.5
                    else:
                        muS -= dmu
                else:
                    break
                if it == 100:
                    raise RuntimeError("Too many")
        self.muS_cache, self.muL_cache = muS, muL
        self.bracket_valid = True
        rnL -= rfill
        rnS -= rfill
        rnc, muc, mud = rnS, muS, 0.0
        flag = False
        for it in range(1, 101):
            if rnc != rnS and rnc != rnL:
                muM = (muL * rnS * rnc * (rnS - rnc) + muS * rnL * rnc * (rnc - rnL)
                       + muc * rnL * rnS * (rnL - rnS)) / ((rnL - rnS) * (rnc - rnL) * (rnc - rnS))
            else:
                muM = muL - rnL * (muL - muS) / (rnL - rnS)
            if ((0.25 * (3.0 * muS + muL) > muM or muM > muL)
                    or (flag and abs(muM - muL) >= abs(muL - muc) * 0.5)
                    or ((not flag) and abs(muM - muL) >= abs(muc - mud) * 0.5)
                    or (flag and abs(muL - muc) < 1.0e-8)
                    or ((not flag) and abs(muc - mud) < 1.0e-8)):
                muM = (muL + muS) * 0.5
                flag = True
            else:
                flag = False
            if abs(muL - muM) < eps:
                break
            mu = muM
            rnM = self.get_rn(muM)
            mud = muc
            muc = muL
            rnc = rnL
            if rnS * (rnM - rfill) < 0:
                muL = muM
                rnL = rnM - rfill
            else:
                muS = muM
                rnS = rnM - rfill
            if abs(rnS) < abs(rnL):
                muL, muS = muS, muL
                rnL, rnS = rnS, rnL
            if it == 100:
                raise RuntimeError("Too many loop!")
        if rnL == rnS:
            mu = (muS + muL) * 0.5
        else:
            mu = (muS * rnL - muL * rnS) / (rnL - rnS)
        self.last_fill = self.get_rn(mu)
        return mu, mu

    def compare_sigma(self, sigf, sig0f, pp):
        """DIIS / linear mixing and esterr (fself.f90:449-542)."""
        m = self.m_diis
        self.i_hist = self.i_hist % m + 1
        self.n_hist = min(self.n_hist + 1, m)
        self.xout[self.i_hist - 1] = sigf.copy()
        self.res[self.i_hist - 1] = sigf - sig0f
        n = self.n_hist
        idx = [((self.i_hist - n + ih - 1) % m) for ih in range(1, n + 1)]
        B = np.zeros((n + 1, n + 1))
        for a, ia in enumerate(idx):
            for b, ib in enumerate(idx):
                B[a, b] = np.sum(np.conj(self.res[ia]) * self.res[ib]).real
            B[a, n] = B[n, a] = 1.0
        for a in range(n):
            B[a, a] += 1e-12
        rhs = np.zeros(n + 1)
        rhs[n] = 1.0
        try:
            c = np.linalg.solve(B, rhs)
        except np.linalg.LinAlgError:
            c = np.zeros(n + 1)
            c[n - 1] = 1.0
        sd = np.zeros_like(sigf)
        for a, ia in enumerate(idx):
            sd = sd + c[a] * self.xout[ia]
        big = np.abs(sd) > 1e-10
        est = np.zeros(sd.shape)
        est[big] = np.abs((sig0f[big] - sd[big]) / sd[big])
        self.esterr = est.max() if big.any() else 0.0
        if n >= 2:
            return sd
        return pp * sd + (1.0 - pp) * sig0f


def mkself(Smat, Cmat, kmap, invk, olist, hamk, eig, uni, mu, fill, temp, Nw, Nx, Ny, Nz,
           sw_out, sw_in, sub_sigma=1, scf_loop=300, eps=1.0e-4, pp=0.3, m_diis=5,
           sw_rescale=False, sw_tail=False, sigma_scale=1.0, sigma_seed=None,
           return_info=False):
    """FLEX self-consistency loop (fself.f90:287-448; wrapper _selfenergy.py:76-127).

    ``sigma_seed=(sigma_py, mu, mu_old)`` plays the role of the sw_in file read.
    Returns (sigmak[No,No,Nw,Nk], mu) like the reference wrapper."""
    Hf = np.asarray(hamk).T
    No = Hf.shape[0]
    Nk = Hf.shape[2]
    Nkall = len(kmap)
    S = np.asarray(Smat).T
    C = np.asarray(Cmat).T
    fl = _Flex(hamk, eig, fill, temp, m_diis, No, Nk, Nw)
    if sw_in:
        sig_py, mu, mu_old = sigma_seed
        sigf = np.asarray(sig_py).T.copy() * sigma_scale
        sig0f = sigf.copy()
        Gf = np.asarray(gen_green(np.ascontiguousarray(sigf.T), hamk, mu, temp)).T
    else:
        mu_old = mu * 1.2
        sigf = np.zeros((Nk, Nw, No, No), dtype=complex)
        sig0f = sigf.copy()
        Gf = np.asarray(gen_Green0(eig, uni, mu, temp, Nw)).T
    info = dict(iters=0, esterr=[], mu=[], stoner=[])
    stoner = 0.0
    for it in range(1, scf_loop + 1):
        Gk_py = np.ascontiguousarray(Gf.T)
        if sw_tail:
            chi, _ = get_chi0_tail(Smat, Cmat, Gk_py, eig, uni, olist, kmap, invk, mu, temp, Nx, Ny, Nz)
        else:
            chi = get_chi0_conv(Gk_py, kmap, invk, olist, temp, Nx, Ny, Nz)
        stoner = ckchi(chi, Smat, Cmat)
        if sw_rescale and stoner >= 1.0:
            chi = chi * ((1.0 - 1.0e-4) / stoner)
        V = get_Vsigma_nosoc_flex(chi, Smat, Cmat)
        sigf = np.asarray(calc_sigma(Gk_py, V, Smat, Cmat, kmap, invk, olist, temp, Nx, Ny, Nz)).T.copy()
        if sub_sigma > 0:
            if sub_sigma == 1:
                base = sigf[:, 0]
            else:
                base = sigf.mean(axis=1)
            sub = np.zeros((Nk, No, No), dtype=complex)
            for l in range(No):
                for m in range(l, No):
                    sub[:, l, m] = (base[:, l, m] + np.conj(base[:, m, l])) * 0.5
                    if l != m:
                        sub[:, m, l] = np.conj(sub[:, l, m])
            sigf = sigf - sub[:, None]
        sigf = fl.compare_sigma(sigf, sig0f, pp)
        info['iters'] = it
        info['esterr'].append(fl.esterr)
        info['stoner'].append(stoner)
        if fl.esterr < eps:
            break
        mu, mu_old = fl.renew_mu(mu, mu_old, sigf)
        info['mu'].append(mu)
        Gf = np.asarray(gen_green(np.ascontiguousarray(sigf.T), hamk, mu, temp)).T
        sig0f = sigf.copy()
    # sigma.bin is written before the final renew_mu (fself.f90:443-446)
    info['mu_file'], info['mu_old_file'] = mu, mu_old
    mu, mu_old = fl.renew_mu(mu, mu_old, sigf)
    out = np.ascontiguousarray(sigf.T)
    if return_info:
        info['mu_old'] = mu_old
        return out, mu, info
    return out, mu


# --------------------------------------------------------------------------- #
#  linearized Eliashberg (feliash.f90)
# --------------------------------------------------------------------------- #
def gap_symms(klist, Norb, gap_sym):
    """Form factors on fractional k (_response.py:415-445)."""
    A = 2 * np.pi
    kx, ky, kz = klist[:, 0], klist[:, 1], klist[:, 2]
    rows = {0: np.ones(len(klist)), 1: np.cos(A * kx) - np.cos(A * ky),
            2: 2 * np.cos(A * kx) * np.cos(A * ky), 3: 2 * np.sin(A * kx) * np.sin(A * ky),
            4: 2 * np.sin(A * kx) * np.sin(A * kz), 5: 2 * np.sin(A * ky) * np.sin(A * kz),
            -1: 2 * np.sin(A * kx), -2: 2 * np.sin(A * ky), -3: 2 * np.sin(A * kx) + 2j * np.sin(A * ky)}
    row = rows.get(gap_sym, np.zeros(len(klist)))
    return np.tile(row, (Norb, 1))


def mkfk_trs_nsoc(Gk, delta):
    """F = -G Delta G^dagger, then Hermitised (feliash.f90:445-484)."""
    Gf = np.asarray(Gk).T
    Df = np.asarray(delta).T
    F = -(Gf @ Df) @ np.conj(np.swapaxes(Gf, -1, -2))
    No = F.shape[-1]
    for l in range(No):
        F[..., l, l] = F[..., l, l].real
        for m in range(l + 1, No):
            F[..., l, m] = np.conj(F[..., m, l])
    return np.ascontiguousarray(F.T)


def _full_anom(Ff, a, b, prt, sgn, dprt, kmap, invk, Nx, Ny, Nz):
    """Full-grid F_(a,b) with parity rules (feliash.f90:566-577)."""
    Nk, Nw = Ff.shape[:2]
    M = 2 * Nw
    out = np.zeros((Nx * Ny * Nz, M), dtype=complex)
    irr = invk[:, 0] - 1
    fl = invk[:, 1]
    sp = _spatial(kmap, Nx, Ny, Nz)
    d = fl == 0
    r = fl == 1
    fab = Ff[:, :, a, b]
    fba = Ff[:, :, b, a]
    pp = prt[a] * prt[b]
    out[sp[d], :Nw] = fab[irr[d]]
    out[sp[d], Nw:] = (sgn * dprt * pp * np.conj(fab[irr[d]]))[:, ::-1]
    out[sp[r], :Nw] = pp * dprt * fab[irr[r]]
    out[sp[r], Nw:] = (sgn * fba[irr[r]])[:, ::-1]
    return out.reshape(Nx, Ny, Nz, M)


def mkdelta_nsoc(fk, Vdelta, Smat, Cmat, kmap, invk, prt, olist, Nx, Ny, Nz, singlet=True):
    """Delta_new = sum_{pairs, l1 >= m2} V_Delta (*) F, Hermitised
    (feliash.f90:486-635); without the T/Nkall weight."""
    Ff = np.asarray(fk).T
    Vf = np.asarray(Vdelta).T
    S = np.asarray(Smat).T
    C = np.asarray(Cmat).T
    ol = olist_f(olist)
    Nk, Nw, No, _ = Ff.shape
    sgn = 1.0 if singlet else -1.0
    dprt = sgn
    nd = np.zeros((Nk, Nw, No, No), dtype=complex)
    for l in range(len(ol)):
        for m in range(len(ol)):
            if ol[l, 0] >= ol[m, 1]:
                v = fft4(_full_boson(Vf, m, l, 0.5 * (S[m, l] + sgn * C[m, l]), kmap, invk, Nx, Ny, Nz))
                f = fft4(_full_anom(Ff, ol[l, 1], ol[m, 0], prt, sgn, dprt, kmap, invk, Nx, Ny, Nz))
                tmp = ifft4(v * f)
                nd[:, :, ol[l, 0], ol[m, 1]] += _gather(tmp, kmap, invk, Nk, Nw)
    for l in range(No):
        for m in range(l, No):
            if l != m:
                nd[:, :, l, m] = np.conj(nd[:, :, m, l])
            else:
                nd[:, :, l, l] = nd[:, :, l, l].real
    return nd


def get_initial_delta(init_delta, uni, kmap, invk, Nw, gap_sym):
    """Delta0 = U diag(d) U^dagger broadcast over w, unit norm with factor 2
    (feliash.f90:637-731).  init_delta Python [Norb,Nk]."""
    d = np.asarray(init_delta)
    No, Nk = d.shape
    Uf = np.asarray(uni).T                       # (No,No,Nk): uni(m,n,k)
    if gap_sym == 0:
        D = np.zeros((Nk, Nw, No, No), dtype=complex)
        for l in range(No):
            D[:, :, l, l] = 1.0
    else:
        U = np.moveaxis(Uf, 2, 0)                # [k][m,n]
        D0 = np.einsum('kmn,nk,kln->kml', U, d, np.conj(U))
        for l in range(No):
            D0[:, l, l] = D0[:, l, l].real
            for m in range(l + 1, No):
                D0[:, l, m] = np.conj(D0[:, m, l])
        D = np.repeat(D0[:, None], Nw, axis=1)
    w = kweight(invk, Nk)
    norm = np.sum(w[:, None, None, None] * np.abs(D) ** 2)
    D = D / np.sqrt(2.0 * norm)
    return np.ascontiguousarray(D.T)


def _arnoldi_two_pass(apply, norm, inner, init, m_dim, info):
    """Two-pass Arnoldi of solve_arnoldi (feliash.f90, feliashwithsoc.f90:190-387):
    pass 1 from the initial gap takes the most negative Ritz value lambda_-;
    unless a Ritz value >= 0.1 appeared, pass 2 on (A + shift) with
    shift = -lambda_-, projected off the pass-1 Ritz vector, takes the largest.
    Returns (unit eigenvector, lambda)."""
    def arnoldi(v0, shift, defl):
        V = [v0]
        H = np.zeros((m_dim + 1, m_dim), dtype=complex)
        m_act = m_dim
        for j in range(m_dim):
            nd = apply(V[j])
            if defl is not None:
                nd = nd + shift * V[j]
                h = inner(defl, nd)
                nd = nd - h * defl
            for ii in range(j + 1):
                h = inner(V[ii], nd)
                H[ii, j] = h
                nd = nd - h * V[ii]
            beta = norm(nd)
            H[j + 1, j] = beta
            if beta < 1e-14:
                m_act = j + 1
                break
            if j < m_dim - 1:
                V.append(nd * (1.0 / beta))
        ev, vr = _geev_right(H[:m_act, :m_act])
        return V, ev, vr, m_act

    V, ev, vr, m_act = arnoldi(init(), 0.0, None)
    idx = int(np.argmin(ev.real))
    lam1 = ev[idx].real
    info['pass1'] = ev
    if ev.real.max() >= 0.1:
        idx = int(np.argmax(ev.real))
        D = sum(vr[ii, idx] * V[ii] for ii in range(m_act))
        return D * (1.0 / norm(D)), ev[idx].real
    d1 = sum(vr[ii, idx] * V[ii] for ii in range(m_act))
    d1 = d1 * (1.0 / norm(d1))
    shift = -lam1
    v0 = init()
    h = inner(d1, v0)
    v0 = v0 - h * d1
    v0 = v0 * (1.0 / norm(v0))
    V, ev, vr, m_act = arnoldi(v0, shift, d1)
    info['pass2'] = ev - shift
    idx = int(np.argmax(ev.real))
    D = sum(vr[ii, idx] * V[ii] for ii in range(m_act))
    return D * (1.0 / norm(D)), ev[idx].real - shift


def linearized_eliashberg(chi, Gk, uni, init_delta, Smat, Cmat, olist, plist, kmap, invk,
                          Nx, Ny, Nz, temp, gap_sym, eps=1.0e-5, itemax=300, arnoldi_m=10,
                          return_info=False):
    """lin_eliash (feliash.f90:1-443; wrapper _eliashberg.py:145-181).
    Returns (delta[No,No,Nw,Nk], lambda)."""
    Gf = np.asarray(Gk).T
    Nk, Nw, No, _ = Gf.shape
    Nkall = len(kmap)
    singlet = gap_sym >= 0
    weight = temp / Nkall
    w = kweight(invk, Nk)[:, None, None, None]
    Vd = get_V_delta(chi, Smat, Cmat, singlet)
    prt = np.asarray(plist, dtype=float)

    def apply(v):
        fk = mkfk_trs_nsoc(Gk, np.ascontiguousarray(v.T))
        return mkdelta_nsoc(fk, Vd, Smat, Cmat, kmap, invk, prt, olist, Nx, Ny, Nz, singlet) * weight

    def norm(f):
        return np.sqrt(2.0 * np.sum(w * np.abs(f) ** 2))

    def inner(u, v):
        return 2.0 * np.sum(w * np.conj(u) * v)

    def init():
        return np.asarray(get_initial_delta(init_delta, uni, kmap, invk, Nw, gap_sym)).T.copy()

    info = {}
    lam_out = 0.0
    if arnoldi_m > 0:
        D, lam_out = _arnoldi_two_pass(apply, norm, inner, init, int(arnoldi_m), info)
    else:
        lam1 = 0.0
        delta1 = None
        D = None
        for i_eig in (1, 2):
            D = init()
            if i_eig == 2:
                proj = inner(delta1, D).real
                D = D - proj * delta1
                D = D * (1.0 / norm(D))
            lam_prev = 0.0
            for it in range(1, itemax + 1):
                nd = apply(D)
                if i_eig == 2:
                    nd = nd - lam1 * D
                    proj = inner(delta1, nd).real
                    nd = nd - proj * delta1
                nrm = norm(nd)
                if nrm <= 0:
                    raise RuntimeError("Error: norm <= 0 in iter_loop")
                inorm = 1.0 / nrm
                lam_rq = inner(D, nd).real
                lam_phys = lam_rq + lam1
                vp = np.sum(w * np.abs(nd * inorm - D) ** 2)
                vn = np.sum(w * np.abs(nd * inorm + D) ** 2)
                verr = np.sqrt(2.0 * min(vp, vn))
                if lam_phys < 0.0 and verr < eps:
                    break
                if verr < eps and abs(lam_phys - lam_prev) < eps * (abs(lam_phys) + 1e-10):
                    break
                lam_prev = lam_phys
                D = nd * inorm
            if i_eig == 1:
                if lam_phys > 0.0:
                    lam_out = lam_phys
                    break
                lam1 = lam_rq
                delta1 = nd * inorm
            else:
                lam_out = lam_phys
        D = nd * (1.0 / norm(nd))
    out = np.ascontiguousarray(D.T)
    if return_info:
        return out, lam_out, info
    return out, lam_out


def _geev_right(A):
    """Eigenvalues and right eigenvectors with LAPACK zgeev normalisation
    (unit 2-norm, largest-modulus component real)."""
    ev, vr = np.linalg.eig(A)
    return ev, vr


def conv_delta_orb_to_band(delta, uni, invk, plist, gap_sym):
    """Band-basis gap at w_0 on the full BZ (feliash.f90:733-785)."""
    Df = np.asarray(delta).T
    Uf = np.asarray(uni).T
    Nkall = len(invk)
    No = Uf.shape[0]
    dprt = -1.0 if gap_sym < 0 else 1.0
    prt = np.asarray(plist, dtype=float)
    out = np.zeros((Nkall, No, No), dtype=complex)
    for i in range(Nkall):
        k = invk[i, 0] - 1
        U = Uf[:, :, k]
        D = Df[k, 0]
        if invk[i, 1] == 0:
            out[i] = np.conj(U.T) @ D @ U
        elif invk[i, 1] == 1:
            L = U.T * prt[None, :]
            t = (L @ D) * (dprt * prt[None, :])
            out[i] = t @ np.conj(U)
    return np.ascontiguousarray(out.T)


def remap_gap(delta0, plist, invk, gap_sym):
    """Irreducible -> full-BZ gap with parity (feliash.f90:787-821)."""
    Df = np.asarray(delta0).T
    prt = np.asarray(plist, dtype=float)
    dprt = -1.0 if gap_sym < 0 else 1.0
    Nkall = len(invk)
    Nk, Nw, No, _ = Df.shape
    out = np.zeros((Nkall, Nw, No, No), dtype=complex)
    for i in range(Nkall):
        k = invk[i, 0] - 1
        if invk[i, 1] == 0:
            out[i] = Df[k]
        elif invk[i, 1] == 1:
            out[i] = dprt * np.outer(prt, prt)[None] * Df[k]
    return np.ascontiguousarray(out.T)


# --------------------------------------------------------------------------- #
#  real-frequency susceptibility (fchi.f90:1-317), Pade continuation (fpade.f90)
# --------------------------------------------------------------------------- #
def get_qshift(klist, qpoint):
    """1-based index of k + q (periodic in [0,1)) for every k of klist[Nk,3]
    (fchi.f90:1-92): coordinates compared on a 2^20 grid; a shifted point off
    the mesh falls back to index 1."""
    scale = 1 << 20
    kl = np.asarray(klist, dtype=np.float64)

    def keys(x):
        i = np.floor(x * scale + 0.5).astype(np.int64) % scale
        return i[:, 0] + i[:, 1] * scale + i[:, 2] * scale * scale
    t = kl + np.asarray(qpoint, dtype=np.float64)[None, :]
    t = np.where(t >= 1.0, t - 1.0, np.where(t < 0.0, t + 1.0, t))
    where = {int(k): j for j, k in enumerate(keys(kl))}
    return np.array([where.get(int(k), 0) + 1 for k in keys(t)], dtype=np.int64)


def _calc_chi(uni, eig, ffermi, olist, temp, qshift, w, idelta, eps):
    """chi0(q, w + i delta) as the Fortran matrix chi(Nchi,Nchi) (calc_chi,
    fchi.f90:98-157): (1/Nk) sum_k sum_{l at k+q, m at k} weight B_i A_j with
    B_i = conj(U_{o_i1,l}(k+q)) U_{o_i2,m}(k), A_j = U_{o_j1,l}(k+q) conj(U_{o_j2,m}(k)),
    weight = (f_l(k+q) - f_m(k)) / (w + e_m(k) - e_l(k+q) + i delta), f (1-f)/T for
    equal energies at w = 0, dropped where |f_l(k+q) - f_m(k)| <= eps."""
    U = np.asarray(uni)                   # [k, band, orb]
    e = np.asarray(eig)
    f = np.asarray(ffermi)
    q0 = np.asarray(qshift) - 1
    Uq, eq, fq = U[q0], e[q0], f[q0]
    dE = e[:, None, :] - eq[:, :, None]   # [k, l, m] = e_m(k) - e_l(k+q)
    df = fq[:, :, None] - f[:, None, :]
    tsafe = max(temp, 1e-12)
    with np.errstate(divide='ignore', invalid='ignore'):
        wt = np.where(np.abs(df) > eps, df / (w + dE + 1j * idelta), 0.0)
    if abs(w) < 1e-12:
        deg = np.abs(dE) < 1e-9
        wt = np.where(deg, (f * (1 - f) / tsafe)[:, None, :] + 0j, wt)
    o1 = np.asarray(olist)[:, 0] - 1
    o2 = np.asarray(olist)[:, 1] - 1
    B = np.conj(Uq[:, :, None, o1]) * U[:, None, :, o2]   # [k, l, m, i]
    A = Uq[:, :, None, o1] * np.conj(U[:, None, :, o2])   # [k, l, m, j]
    return np.einsum('klm,klmi,klmj->ij', wt, B, A) / len(e)


def get_chi_irr(uni, eig, ffermi, qshift, olist, wlist, idelta, temp):
    """[Nw, Nchi, Nchi] (Python view of the Fortran chi(Nchi,Nchi,Nw));
    fchi.f90:206-240, wrapper _susceptibility.py:271-293 (eps = idelta 1e-3)."""
    eps = idelta * 1e-3
    return np.array([_calc_chi(uni, eig, ffermi, olist, temp, qshift, w, idelta, eps).T for w in wlist])


def chis_qmap(uni, eig, ffermi, klist, Smat, olist, Nx, Ny, temp, ecut, idelta):
    """(chis, chi) traces over the orbital-diagonal pairs on the (qx, qy)
    plane at w = ecut, [Nx, Ny] each (fchi.f90:242-317, wrapper
    _susceptibility.py:295-324): chi_s = (I - chi0 S)^-1 chi0 per q."""
    eps = idelta * 1e-3
    S = np.asarray(Smat).T
    ol = np.asarray(olist)
    diag = np.nonzero(ol[:, 0] == ol[:, 1])[0]
    chis = np.zeros((Nx, Ny), dtype=complex)
    chi = np.zeros((Nx, Ny), dtype=complex)
    for i in range(Nx):
        for j in range(Ny):
            qs = get_qshift(klist, [i / Nx, j / Ny, 0.0])
            c = _calc_chi(uni, eig, ffermi, olist, temp, qs, ecut, idelta, eps)
            cs = np.linalg.solve(np.eye(len(S)) - c @ S, c)
            chis[i, j] = cs[np.ix_(diag, diag)].sum()
            chi[i, j] = c[np.ix_(diag, diag)].sum()
    return chis, chi


def get_tr_chi(chis, chi0, olist):
    """(trchis[Nw], trchi0[Nw], chis_orb[Nw, Norb+2]) (fchi.f90:160-204)."""
    S = np.asarray(chis).T                 # Fortran chis(k, j, w)
    C = np.asarray(chi0).T
    ol = np.asarray(olist)
    Nchi, Nw = S.shape[0], S.shape[2]
    Norb = int(ol.max())
    ts = np.zeros(Nw, dtype=complex)
    t0 = np.zeros(Nw, dtype=complex)
    so = np.zeros((Norb + 2, Nw), dtype=complex)
    for i in range(Nw):
        for j in range(Nchi):
            if ol[j, 0] != ol[j, 1]:
                continue
            for k in range(Nchi):
                if ol[k, 0] != ol[k, 1]:
                    continue
                ts[i] += S[k, j, i]
                t0[i] += C[k, j, i]
                if ol[j, 0] == ol[k, 0]:
                    so[ol[j, 0] - 1, i] = S[k, j, i]
                if ol[k, 0] == 2 and ol[j, 0] == 3:
                    so[Norb, i] = S[k, j, i]
                if ol[k, 0] == 2 and ol[j, 0] == 4:
                    so[Norb + 1, i] = S[k, j, i]
    return ts, t0, np.ascontiguousarray(so.T)


def get_a(inp_data, xlist):
    """Continued-fraction (Thiele) coefficients of the data at the points
    (fpade.f90:1-42): g_1 = data, g_i(z_j) = (g_{i-1}(z_{i-1}) - g_{i-1}(z_j))
    / ((z_j - z_{i-1}) g_{i-1}(z_j)) (0 where g_{i-1}(z_j) = 0), a_i = g_i(z_i)."""
    g0 = np.asarray(inp_data, dtype=complex).copy()
    x = np.asarray(xlist, dtype=complex)
    Np = len(g0)
    a = np.zeros(Np, dtype=complex)
    a[0] = g0[0]
    with np.errstate(divide='ignore', invalid='ignore'):
        for i in range(1, Np):
            z = g0 == 0
            den = np.where(z, 1.0, (x - x[i - 1]) * g0)
            g0 = np.where(z, 0.0, (g0[i - 1] - g0) / den)
            a[i] = g0[i]
    return a


def get_QP(a, xlist, wlist):
    """(Q, P): denominator and numerator of the continued fraction at wlist
    (fpade.f90:44-85; the wrapper returns them in this order)."""
    a = np.asarray(a, dtype=complex)
    x = np.asarray(xlist, dtype=complex)
    z = np.asarray(wlist, dtype=complex)
    P0, P1 = np.zeros_like(z), np.full_like(z, a[0])
    Q0, Q1 = np.ones_like(z), np.ones_like(z)
    P, Q = P1.copy(), Q1.copy()
    for i in range(1, len(a)):
        fct = a[i] * (z - x[i - 1])
        P = P1 + fct * P0
        Q = Q1 + fct * Q0
        P0, P1, Q0, Q1 = P1, P, Q1, Q
    return Q, P


def pade_with_trace(A, iwlist, wlist):
    """[Nk, Nw]: sum over the orbital-diagonal blocks of A[Norb,Norb,Niw,Nk]
    of their Pade continuations (fpade.f90:87-129)."""
    A = np.asarray(A)
    Norb, Nk = A.shape[0], A.shape[3]
    out = np.zeros((Nk, len(wlist)), dtype=complex)
    for o in range(Norb):
        for k in range(Nk):
            a = get_a(A[o, o, :, k], iwlist)
            Q, P = get_QP(a, iwlist, wlist)
            out[k] += P / Q
    return out


# --------------------------------------------------------------------------- #
#  sigma.bin (fself.f90:742-780): Fortran unformatted sequential records
# --------------------------------------------------------------------------- #
def write_sigma_bin(path, mu, mu_old, sigmak, temp, Nw, Norb, Nk):
    import struct
    def rec(b):
        return struct.pack('<i', len(b)) + b + struct.pack('<i', len(b))
    with open(path, 'wb') as f:
        f.write(rec(struct.pack('<d', mu)))
        f.write(rec(struct.pack('<d', mu_old)))
        f.write(rec(np.ascontiguousarray(sigmak, dtype=np.complex128).tobytes()))
        f.write(rec(struct.pack('<4d', temp, float(Nw), float(Norb), float(Nk))))


# --------------------------------------------------------------------------- #
#  Wannier-R output (libs/plibs/_wannier_io.py)
# --------------------------------------------------------------------------- #
def wannier_r(data, kmap, invk, Nx, Ny, Nz, N_cut, full=False):
    """data[No,No,Nw,Nk] -> data_R[No,No,min(N_cut,Nw),Nx,Ny,Nz].

    The irreducible values are placed on the grid through kmap, conjugated
    where invk flags the time-reversed partner (_wannier_io.py:48-68; full:
    data already on every grid point in kmap order, :195-199), then numpy's
    normalised inverse FFT over the three k axes, first N_cut frequencies
    (:81-83)."""
    data = np.asarray(data)
    No, _, Nw, _ = data.shape
    grid = np.zeros((No, No, Nw, Nx, Ny, Nz), dtype=np.complex128)
    ix, iy, iz = kmap[:, 0], kmap[:, 1], kmap[:, 2]
    if full:
        grid[:, :, :, ix, iy, iz] = data
    else:
        v = data[..., invk[:, 0] - 1]
        tr = invk[:, 1] != 0
        v[..., tr] = np.conj(v[..., tr])
        grid[:, :, :, ix, iy, iz] = v
    return np.fft.ifftn(grid, axes=(-3, -2, -1))[:, :, :min(N_cut, Nw)]


def wannier_select(data_R, Nx, Ny, Nz, zero_tol):
    """Zero-tolerance cut and nonzero-R selection of _wannier_io.py:84-99
    (on a copy): (data_out[No,No,Ncut,Nr], rvec[Nr,3])."""
    d = np.array(data_R)
    tol = np.abs(d).max() * zero_tol
    d.real[np.abs(d.real) < tol] = 0.0
    d.imag[np.abs(d.imag) < tol] = 0.0
    sel = np.abs(d).max(axis=(0, 1, 2)) > 0
    ix, iy, iz = np.where(sel)
    rv = np.stack([np.where(ix <= Nx // 2, ix, ix - Nx), np.where(iy <= Ny // 2, iy, iy - Ny),
                   np.where(iz <= Nz // 2, iz, iz - Nz)], axis=1).astype(np.int64)
    return d[:, :, :, ix, iy, iz], rv


# --------------------------------------------------------------------------- #
#  spin-orbit coupling (fchi0withsoc.f90, fselfwithsoc.f90, fUJ.f90:109-251,
#  367-399).  Orbitals carry spin: slist[a] = +-1, invs[a] (1-based) is the
#  spin-flipped partner of a; olist runs over all Norb^2 pairs.
# --------------------------------------------------------------------------- #
def soc_spin_lists(Norb):
    """Default slist / invs of the reference driver (main.py:1373-1385): the
    first (Norb+1)//2 orbitals spin up, the rest down, partners in order."""
    slist = np.ones(Norb, dtype=np.int64)
    slist[(Norb + 1) // 2:] = -1
    invs = np.concatenate([np.arange((Norb + 1) // 2, Norb), np.arange((Norb + 1) // 2)]) + 1
    return slist, invs.astype(np.int64)


def get_vmat_soc(olist, slist, site, invs, U, J):
    """SOC interaction vertex (fUJ.f90:109-180).  Python V[i,j] is Fortran
    Vmat(j,i), the value the reference assigns for the pair (i, j)."""
    ol = olist_f(olist)
    sl = np.asarray(slist)
    inv = np.asarray(invs) - 1
    Nchi = len(ol)
    V = np.zeros((Nchi, Nchi))
    Up = U - 2 * J
    for i in range(Nchi):
        for j in range(Nchi):
            if site[i] != site[j]:
                continue
            a, b = ol[i]
            c, d = ol[j]
            if sl[a] == sl[b] and sl[c] == sl[d]:
                if sl[a] == sl[c]:                       # same-spin density channel
                    if a == d and c == b:
                        continue
                    if a == b and c == d:
                        V[i, j] = Up - J
                    elif a == c and b == d:
                        V[i, j] = J - Up
                else:                                    # opposite-spin density channel
                    if a == b and c == d:
                        V[i, j] = U if a == inv[c] else Up
                    elif a == inv[c] and b == inv[d]:
                        V[i, j] = J
                    elif a == inv[d] and c == inv[b]:
                        V[i, j] = J
            elif sl[a] == sl[c] and sl[b] == sl[d]:      # spin-flip channel
                if a == inv[b] and c == inv[d]:
                    V[i, j] = -U if a == c else -J
                elif a == c and b == d:
                    V[i, j] = -Up
                elif a == inv[d] and c == inv[b]:
                    V[i, j] = -J
    return V


def get_chi_map_soc(olist, invs):
    """Exchange-symmetry partners chi_{1234} = sgn chi*_{-4-3-2-1} with spin
    flips (fchi0withsoc.f90:216-279): cmap[m1,l1] = (m2,l2), irr[i] = (l1,m1),
    0-based."""
    ol = olist_f(olist)
    inv = np.asarray(invs) - 1
    Nchi = len(ol)
    cmap = np.zeros((Nchi, Nchi, 2), dtype=np.int64)
    cirr = np.zeros((Nchi, Nchi), dtype=np.int64)
    for l1 in range(Nchi):
        for m1 in range(Nchi):
            t1 = (ol[l1, 0], ol[l1, 1], ol[m1, 0], ol[m1, 1])
            found = False
            for l2 in range(Nchi):
                for m2 in range(Nchi):
                    if t1 == (inv[ol[m2, 1]], inv[ol[m2, 0]], inv[ol[l2, 1]], inv[ol[l2, 0]]):
                        cmap[m1, l1] = (m2, l2)
                        if cirr[m2, l2] == 0:
                            cirr[m1, l1] = 1
                        found = True
                        break
                if found:
                    break
    irr = [(l1, m1) for l1 in range(Nchi) for m1 in range(Nchi) if cirr[m1, l1] == 1]
    return cmap, np.array(irr, dtype=np.int64).reshape(-1, 2)


def soc_signs(olist, slist):
    """sgnsig(i,j) = s_i s_j and sgnsig2(l,m) = s_l1 s_l2 s_m1 s_m2
    (fchi0withsoc.f90:136-154)."""
    ol = olist_f(olist)
    s = np.asarray(slist, dtype=float)
    sg = np.outer(s, s)
    p = s[ol[:, 0]] * s[ol[:, 1]]
    return sg, np.outer(p, p)


def get_invschi(olist, invs):
    """invschi(l) = 1-based index of the spin-flipped pair of l, or 0
    (fchi0withsoc.f90:200-213)."""
    ol = olist_f(olist)
    inv = np.asarray(invs) - 1
    out = np.zeros(len(ol), dtype=np.int64)
    for l in range(len(ol)):
        hit = np.nonzero((ol[:, 0] == inv[ol[l, 0]]) & (ol[:, 1] == inv[ol[l, 1]]))[0]
        if len(hit):
            out[l] = hit[0] + 1
    return out


def _full_green_soc(Gf, a, b, sg, inv, kmap, invk, Nx, Ny, Nz):
    """Full-grid G_ab(k, slot) with the time-reversal rule of a spinful
    basis (fchi0withsoc.f90:57-75): G_ab(-k, w) = s_a s_b G_{-b,-a}(k, w),
    -w through G(k,-w) = G(k,w)^dagger."""
    Nk, Nw = Gf.shape[:2]
    M = 2 * Nw
    out = np.zeros((Nx * Ny * Nz, M), dtype=complex)
    irr = invk[:, 0] - 1
    fl = invk[:, 1]
    sp = _spatial(kmap, Nx, Ny, Nz)
    d = fl == 0
    r = fl == 1
    out[sp[d], :Nw] = Gf[irr[d], :, a, b]
    out[sp[d], Nw:] = np.conj(Gf[irr[d], :, b, a])[:, ::-1]
    out[sp[r], :Nw] = sg[b, a] * Gf[irr[r], :, inv[b], inv[a]]
    out[sp[r], Nw:] = (sg[a, b] * np.conj(Gf[irr[r], :, inv[a], inv[b]]))[:, ::-1]
    return out.reshape(Nx, Ny, Nz, M)


def get_chi0_conv_soc(Gk, kmap, invk, invs, olist, slist, temp, Nx, Ny, Nz):
    """chi0 of a spinful basis by FFT convolution (fchi0withsoc.f90:1-109).
    Returns chi[Nchi,Nchi,Nw,Nk]."""
    Gf = np.asarray(Gk).T
    Nk, Nw = Gf.shape[:2]
    ol = olist_f(olist)
    inv = np.asarray(invs) - 1
    sg, sg2 = soc_signs(olist, slist)
    cmap, irr = get_chi_map_soc(olist, invs)
    Nchi = len(ol)
    weight = temp / len(kmap)
    chif = np.zeros((Nk, Nw, Nchi, Nchi), dtype=complex)
    for l, m in irr:
        g13 = fft4(_full_green_soc(Gf, ol[l, 0], ol[m, 0], sg, inv, kmap, invk, Nx, Ny, Nz))
        g42 = fft4(_full_green_soc(Gf, ol[m, 1], ol[l, 1], sg, inv, kmap, invk, Nx, Ny, Nz))
        tmp = ifft4(-g13 * _reverse(g42))
        chif[:, :, m, l] = _threshold(_gather(tmp, kmap, invk, Nk, Nw) * weight)
        pm, pl = cmap[m, l]
        chif[:, :, pm, pl] = sg2[m, l] * np.conj(chif[:, :, m, l])
    return np.ascontiguousarray(chif.T)


def ckchi_soc(chi, Vmat):
    """max_q max Re eig(-chi0(q,0) V) (fselfwithsoc.f90:191-226)."""
    chif = np.asarray(chi).T
    Vf = np.asarray(Vmat).T
    return max(np.linalg.eigvals(-chif[i, 0] @ Vf).real.max() for i in range(chif.shape[0]))


def get_chi0_soc(Vmat, Gk, olist, slist, kmap, invk, invs, temp, Nx, Ny, Nz):
    """(chi, sgnsig, sgnsig2, invschi) as the reference wrapper returns them
    (fchi0withsoc.f90:111-214, _susceptibility.py:496-528)."""
    chi = get_chi0_conv_soc(Gk, kmap, invk, invs, olist, slist, temp, Nx, Ny, Nz)
    sg, sg2 = soc_signs(olist, slist)
    return chi, sg, sg2, get_invschi(olist, invs)


def get_V_soc_flex(chi, Vmat):
    """-V + V (I + chi0 V)^-1 chi0 V per (q, i nu) (fUJ.f90:367-399)."""
    chif = np.asarray(chi).T
    Vf = np.asarray(Vmat).T
    AV = chif @ Vf
    X = np.linalg.solve(np.eye(len(Vf)) + AV, AV)
    return np.ascontiguousarray((-Vf + Vf @ X).T)


def calc_sigma_soc(Gk, Vsigma, Vmat, kmap, invk, invs, olist, slist, temp, Nx, Ny, Nz):
    """Sigma = (T/Nkall) sum over all Nchi^2 pairs of V (*) G in a spinful
    basis, the nu -> infinity slot -Vmat (fselfwithsoc.f90:540-667)."""
    Gf = np.asarray(Gk).T
    Vf = np.asarray(Vsigma).T
    Vm = np.asarray(Vmat).T
    ol = olist_f(olist)
    inv = np.asarray(invs) - 1
    sg, _ = soc_signs(olist, slist)
    Nk, Nw, No, _ = Gf.shape
    Nchi = len(ol)
    sig = np.zeros((Nk, Nw, No, No), dtype=complex)
    gft = {}
    for l in range(Nchi):
        for m in range(Nchi):
            v = fft4(_full_boson(Vf, m, l, -Vm[m, l], kmap, invk, Nx, Ny, Nz))
            key = (ol[m, 1], ol[l, 1])
            if key not in gft:
                gft[key] = fft4(_full_green_soc(Gf, key[0], key[1], sg, inv, kmap, invk, Nx, Ny, Nz))
            sig[:, :, ol[m, 0], ol[l, 0]] += _gather(ifft4(v * gft[key]), kmap, invk, Nk, Nw)
    sig *= temp / len(kmap)
    return np.ascontiguousarray(_threshold(sig).T)


def calc_sigma_soc_direct(Gk, Vsigma, Vmat, kmap, invk, invs, olist, slist, temp, Nx, Ny, Nz):
    """The same Sigma as a direct circular sum over (q, nu) (independent check)."""
    Gf = np.asarray(Gk).T
    Vf = np.asarray(Vsigma).T
    Vm = np.asarray(Vmat).T
    ol = olist_f(olist)
    inv = np.asarray(invs) - 1
    sg, _ = soc_signs(olist, slist)
    Nk, Nw, No, _ = Gf.shape
    M = 2 * Nw
    sig = np.zeros((Nk, Nw, No, No), dtype=complex)
    d = np.nonzero(invk[:, 1] == 0)[0]
    for l in range(len(ol)):
        for m in range(len(ol)):
            v = _full_boson(Vf, m, l, -Vm[m, l], kmap, invk, Nx, Ny, Nz)
            g = _full_green_soc(Gf, ol[m, 1], ol[l, 1], sg, inv, kmap, invk, Nx, Ny, Nz)
            for i in d:
                kx, ky, kz = kmap[i]
                gs = g[(kx - np.arange(Nx)) % Nx][:, (ky - np.arange(Ny)) % Ny][:, :, (kz - np.arange(Nz)) % Nz]
                for n in range(Nw):
                    sig[invk[i, 0] - 1, n, ol[m, 0], ol[l, 0]] += (v * gs[..., (n - np.arange(M)) % M]).sum()
    sig *= temp / len(kmap)
    return np.ascontiguousarray(sig.T)


def get_chis_chic_soc(chi, Vmat, olist, slist, invs):
    """Charge, longitudinal and transverse spin susceptibilities at the first
    boson frequency (fchi0withsoc.f90:281-326; orb_list of
    _susceptibility.py:545-556).  Returns (chic, chiszz, chispm)
    [Nchi/4, Nchi/4, Nk]."""
    chif = np.asarray(chi).T
    Vf = np.asarray(Vmat).T
    ol = olist_f(olist)
    sl = np.asarray(slist)
    inv = np.asarray(invs) - 1
    Nk = chif.shape[0]
    Nchi = len(ol)
    up = np.nonzero(sl == 1)[0]
    nu = len(up)
    # pair n = i nu + j of the spin-up orbitals is (up[j], up[i])
    pos = {(up[j], up[i]): i * nu + j for i in range(nu) for j in range(nu)}
    orb = np.zeros(Nchi, dtype=np.int64)
    for i in range(Nchi):
        a = ol[i, 0] if sl[ol[i, 0]] == 1 else inv[ol[i, 0]]
        b = ol[i, 1] if sl[ol[i, 1]] == 1 else inv[ol[i, 1]]
        orb[i] = pos[(a, b)]
    n4 = Nchi // 4
    chic = np.zeros((Nk, n4, n4), dtype=complex)
    szz = np.zeros((Nk, n4, n4), dtype=complex)
    spm = np.zeros((Nk, n4, n4), dtype=complex)
    s1, s2 = sl[ol[:, 0]], sl[ol[:, 1]]
    for q in range(Nk):
        A = chif[q, 0]
        X = np.linalg.solve(np.eye(Nchi) + A @ Vf, A)
        for l in range(Nchi):
            for m in range(Nchi):
                if s1[l] == s1[m] and s2[l] == s2[m] and s1[l] == -s2[l] and s1[l] == 1:
                    spm[q, orb[l], orb[m]] += X[m, l]
                if s1[l] == s2[l] and s1[m] == s2[m]:
                    chic[q, orb[l], orb[m]] += X[m, l]
                    szz[q, orb[l], orb[m]] += X[m, l] if s1[l] == s1[m] else -X[m, l]
    return (np.ascontiguousarray((0.5 * chic).T), np.ascontiguousarray((0.25 * szz).T),
            np.ascontiguousarray(spm.T))


def mkself_soc(Vmat, kmap, invk, invs, olist, slist, hamk, eig, uni, mu, fill, temp, Nw, Nx, Ny, Nz,
               sw_out, sw_in, sub_sigma=1, scf_loop=300, eps=1.0e-4, pp=0.3, m_diis=5, sw_rescale=False,
               sigma_scale=1.0, sigma_seed=None, return_info=False):
    """FLEX loop in a spinful basis (fselfwithsoc.f90:1-539; wrapper
    _selfenergy.py:129-174): the loop of mkself with the spinful chi0, the
    Stoner factor of -chi0 V, the vertex -V + V chi_RPA V and no tail."""
    Hf = np.asarray(hamk).T
    No, Nk = Hf.shape[0], Hf.shape[2]
    fl = _Flex(hamk, eig, fill, temp, m_diis, No, Nk, Nw)
    if sw_in:
        sig_py, mu, mu_old = sigma_seed
        sigf = np.asarray(sig_py).T.copy() * sigma_scale
        sig0f = sigf.copy()
        Gf = np.asarray(gen_green(np.ascontiguousarray(sigf.T), hamk, mu, temp)).T
    else:
        mu_old = mu * 1.2
        sigf = np.zeros((Nk, Nw, No, No), dtype=complex)
        sig0f = sigf.copy()
        Gf = np.asarray(gen_Green0(eig, uni, mu, temp, Nw)).T
    info = dict(iters=0, esterr=[], mu=[], stoner=[])
    for it in range(1, scf_loop + 1):
        Gk_py = np.ascontiguousarray(Gf.T)
        chi = get_chi0_conv_soc(Gk_py, kmap, invk, invs, olist, slist, temp, Nx, Ny, Nz)
        stoner = ckchi_soc(chi, Vmat)
        if sw_rescale and stoner >= 1.0:
            chi = chi * ((1.0 - 1.0e-4) / stoner)
        V = get_V_soc_flex(chi, Vmat)
        sigf = np.asarray(calc_sigma_soc(Gk_py, V, Vmat, kmap, invk, invs, olist, slist, temp,
                                         Nx, Ny, Nz)).T.copy()
        if sub_sigma > 0:
            base = sigf[:, 0] if sub_sigma == 1 else sigf.mean(axis=1)
            sub = 0.5 * (base + np.conj(np.swapaxes(base, 1, 2)))
            sigf = sigf - sub[:, None]
        sigf = fl.compare_sigma(sigf, sig0f, pp)
        info['iters'] = it
        info['esterr'].append(fl.esterr)
        info['stoner'].append(stoner)
        if fl.esterr < eps:
            break
        mu, mu_old = fl.renew_mu(mu, mu_old, sigf)
        info['mu'].append(mu)
        Gf = np.asarray(gen_green(np.ascontiguousarray(sigf.T), hamk, mu, temp)).T
        sig0f = sigf.copy()
    info['mu_file'], info['mu_old_file'] = mu, mu_old
    mu, mu_old = fl.renew_mu(mu, mu_old, sigf)
    out = np.ascontiguousarray(sigf.T)
    if return_info:
        return out, mu, info
    return out, mu


# --------------------------------------------------------------------------- #
#  linearized Eliashberg in a spinful basis (feliashwithsoc.f90).  The gap
#  lives on the full zone: Python [No, No, Nw, Nkall], Fortran view [i, j, a, b].
# --------------------------------------------------------------------------- #
def get_initial_delta_soc(init_delta, uni, kmap, invk, slist, invs, Nw, gap_sym):
    """feliashwithsoc.f90:665-748: flag-0 points -U diag(d) (s_b conj U(-b,:))^T,
    flag-1 points (s_a conj U(-a,:)) diag(d) U^T, the same at every frequency,
    same-spin blocks zero for gap_sym > 0, normalised to sqrt(2 sum |.|^2) = 1."""
    Uf = np.asarray(uni).T                    # uni(a, n, k)
    df0 = np.asarray(init_delta).T            # init(k, n)
    sl = np.asarray(slist, dtype=float)
    inv = np.asarray(invs) - 1
    Nkall = len(invk)
    No = Uf.shape[0]
    D = np.zeros((Nkall, No, No), dtype=complex)
    for i in range(Nkall):
        k, f = invk[i, 0] - 1, invk[i, 1]
        U = Uf[:, :, k]
        d = df0[k]
        if f == 0:
            D[i] = -(U * d[None, :]) @ (sl[:, None] * np.conj(U[inv, :])).T
        elif f == 1:
            D[i] = (sl[:, None] * np.conj(U[inv, :]) * d[None, :]) @ U.T
    if gap_sym > 0:
        D[:, np.outer(sl, sl) >= 0] = 0.0
    full = np.repeat(D[:, None], Nw, axis=1)  # [i, j, a, b]
    full = full / np.sqrt(2.0 * np.sum(np.abs(full) ** 2))
    return np.ascontiguousarray(full.T)


def mkfk_trs_soc(Gk, delta, sgnsig, slist, invk, invs, gap_sym):
    """F = -G(k) Delta(k) G(-k)^dagger on the full zone with its spin structure
    (feliashwithsoc.f90:426-518).  Returns Python [No, No, Nw, Nkall]."""
    Gf = np.asarray(Gk).T                     # [k, j, a, b]
    Df = np.asarray(delta).T                  # [i, j, a, b]
    sg = np.asarray(sgnsig).T                 # sgnsig(a, b)
    sl = np.asarray(slist)
    inv = np.asarray(invs) - 1
    Nkall, Nw, No, _ = Df.shape
    F = np.zeros_like(Df)
    for i in range(Nkall):
        k = invk[i, 0] - 1
        G = Gf[k]                              # [j, a, b]
        if invk[i, 1] == 0:
            g1 = G
            g2 = sg.T[None] * G[:, inv][:, :, inv]          # g2(l, n) = sg(n, l) G(-l, -n)
        else:
            g1 = sg.T[None] * np.swapaxes(G[:, inv][:, :, inv], 1, 2)   # g1(l, n) = sg(n, l) G(-n, -l)
            g2 = np.swapaxes(G, 1, 2)                        # g2(l, n) = G(n, l)
        F[i] = -(g1 @ Df[i]) @ np.conj(np.swapaxes(g2, 1, 2))
    for l in range(No):
        if sl[l] != 1:
            continue
        for m in range(No):
            if sl[m] == 1:
                if gap_sym < 0:
                    F[:, :, inv[l], inv[m]] = -np.conj(F[:, :, m, l])
                else:
                    F[:, :, m, l] = 0.0
                    F[:, :, inv[l], inv[m]] = 0.0
            elif inv[m] == l:
                F[:, :, inv[l], inv[m]] = F[:, :, m, l].real
    for l in range(No):
        if sl[l] != 1:
            continue
        for m in range(No):
            if sl[m] != -1 or gap_sym == -10:
                continue
            F[:, :, inv[m], inv[l]] = F[:, :, m, l] if gap_sym < 0 else -F[:, :, m, l]
    return np.ascontiguousarray(F.T)


def _full_boson_soc(Vf, m, l, li, mi, sg2, static, kmap, invk, Nx, Ny, Nz):
    """V_(m,l)(q, slot) of a spinful basis (feliashwithsoc.f90:545-573): -k
    points read sg2 V(q)_(l',m') at +nu and sg2 conj V(q)_(m',l') at -nu."""
    Nk, Nw = Vf.shape[:2]
    M = 2 * Nw
    out = np.zeros((Nx * Ny * Nz, M), dtype=complex)
    irr = invk[:, 0] - 1
    fl = invk[:, 1]
    sp = _spatial(kmap, Nx, Ny, Nz)
    d, r = fl == 0, fl == 1
    out[sp[d], :Nw] = Vf[irr[d], :, m, l]
    out[sp[r], :Nw] = sg2 * Vf[irr[r], :, li, mi]
    out[:, Nw] = static
    out[sp[d], Nw + 1:] = np.conj(Vf[irr[d], 1:, l, m])[:, ::-1]
    out[sp[r], Nw + 1:] = (sg2 * np.conj(Vf[irr[r], 1:, mi, li]))[:, ::-1]
    return out.reshape(Nx, Ny, Nz, M)


def _soc_gap_pairs(ol, sl, inv, gap_sym):
    """(l, m) pairs mkdelta_soc convolves (feliashwithsoc.f90:538-540)."""
    out = []
    for l in range(len(ol)):
        for m in range(len(ol)):
            l1, m2 = ol[l, 0], ol[m, 1]
            ss = sl[l1] * sl[m2]
            if not (gap_sym < 0 or (gap_sym > 0 and ss < 0)):
                continue
            if not ((gap_sym == -10 and ss < 0) or sl[l1] == 1):
                continue
            if not (sl[m2] == 1 or l1 >= inv[m2]):
                continue
            out.append((l, m))
    return out


def mkdelta_soc(fk, Vdelta, Vmat, kmap, invk, invs, invschi, olist, slist, Nx, Ny, Nz, gap_sym):
    """newdelta = sum over the pairs of V_Delta (*) F on the full zone, then its
    spin structure (feliashwithsoc.f90:520-663); un-weighted.  Python in / out
    [No, No, Nw, Nkall]."""
    Ff = np.asarray(fk).T                     # [i, j, a, b]
    Vf = np.asarray(Vdelta).T
    Vm = np.asarray(Vmat).T
    ol = olist_f(olist)
    sl = np.asarray(slist)
    inv = np.asarray(invs) - 1
    isc = np.asarray(invschi) - 1
    _, sg2 = soc_signs(olist, slist)
    Nkall, Nw, No, _ = Ff.shape
    M = 2 * Nw
    sp = _spatial(kmap, Nx, Ny, Nz)
    minus = invk[:, 2] - 1
    nd = np.zeros_like(Ff)
    for l, m in _soc_gap_pairs(ol, sl, inv, gap_sym):
        v = fft4(_full_boson_soc(Vf, m, l, isc[l], isc[m], sg2[m, l], -Vm[m, l], kmap, invk, Nx, Ny, Nz))
        g = np.zeros((Nx * Ny * Nz, M), dtype=complex)
        g[sp, :Nw] = Ff[:, :, ol[l, 1], ol[m, 0]]
        g[sp, Nw:] = -Ff[minus, :, ol[m, 0], ol[l, 1]][:, ::-1]
        tmp = ifft4(v * fft4(g.reshape(Nx, Ny, Nz, M))).reshape(-1, M)
        nd[:, :, ol[l, 0], ol[m, 1]] += tmp[sp, :Nw]
    for l in range(No):
        if sl[l] != 1:
            continue
        for m in range(No):
            if sl[m] == 1:
                nd[:, :, inv[l], inv[m]] = -np.conj(nd[:, :, m, l])
            elif inv[m] == l:
                nd[:, :, l, m] = nd[:, :, l, m].real
            elif l > inv[m]:
                nd[:, :, inv[m], inv[l]] = np.conj(nd[:, :, l, m])
    for l in range(No):
        if sl[l] != 1:
            continue
        for m in range(No):
            if sl[m] != -1 or gap_sym == -10:
                continue
            nd[:, :, inv[l], inv[m]] = -nd[:, :, l, m] if gap_sym > 0 else nd[:, :, l, m]
    return np.ascontiguousarray(nd.T)


def linearized_eliashberg_soc(chi, Gk, uni, init_delta, Vmat, sgnsig, sgnsig2, plist, slist, olist, kmap, invk,
                              invs, invschi, Nx, Ny, Nz, temp, gap_sym, eps=1.0e-4, itemax=300, arnoldi_m=10,
                              return_info=False):
    """lin_eliash_soc (feliashwithsoc.f90:1-424; wrapper _eliashberg.py:183-223).
    Returns (delta[No,No,Nw,Nkall], lambda); the reference wrapper returns
    delta and prints lambda."""
    weight = temp / len(kmap)
    Vd = get_V_soc_flex(chi, Vmat)
    Nw = np.asarray(Gk).shape[2]

    def apply(v):
        fk = mkfk_trs_soc(Gk, np.ascontiguousarray(v.T), sgnsig, slist, invk, invs, gap_sym)
        return -np.asarray(mkdelta_soc(fk, Vd, Vmat, kmap, invk, invs, invschi, olist, slist, Nx, Ny, Nz,
                                       gap_sym)).T * weight

    def norm(f):
        return np.sqrt(2.0 * np.sum(np.abs(f) ** 2))

    def inner(u, v):
        return 2.0 * np.sum(np.conj(u) * v)

    def init():
        return np.asarray(get_initial_delta_soc(init_delta, uni, kmap, invk, slist, invs, Nw, gap_sym)).T.copy()

    info = {}
    if arnoldi_m > 0:
        D, lam = _arnoldi_two_pass(apply, norm, inner, init, int(arnoldi_m), info)
    else:
        norm2 = 0.0
        for i_eig in (1, 2):
            D = init()
            count = 0
            for it in range(1, itemax + 1):
                nd = apply(D) + D * norm2
                nrm = norm(nd)
                inorm = 1.0 / nrm
                lam_rq = inner(D, nd).real
                vp = np.sum(np.abs(nd * inorm - D) ** 2)
                vn = np.sum(np.abs(nd * inorm + D) ** 2)
                verr = np.sqrt(2.0 * min(vp, vn))
                if verr < eps:
                    if lam_rq - norm2 > 0.1 or lam_rq - norm2 < 0.0:
                        break
                    if verr < eps * 1e-2:
                        count += 1
                    if count > 30:
                        break
                D = nd * inorm
            lam = lam_rq - norm2
            if i_eig == 1 and lam > 0.0:
                break
            norm2 = nrm
        D = nd * (1.0 / norm(nd))
    out = np.ascontiguousarray(D.T)
    if return_info:
        return out, lam, info
    return out, lam


def conv_delta_orb_to_band_soc(delta, uni, invk, invs, slist):
    """Band-basis gap at the first frequency (feliashwithsoc.f90:750-794):
    flag 0: U^dagger Delta (s_n U(-n,:)), flag 1: (-s_n U(-n,:))^T-column form
    Delta conj(U).  Returns Python [No, No, Nkall]."""
    Df = np.asarray(delta).T                  # [i, j, a, b]
    Uf = np.asarray(uni).T                    # uni(a, n, k)
    sl = np.asarray(slist, dtype=float)
    inv = np.asarray(invs) - 1
    Nkall, _, No, _ = Df.shape
    out = np.zeros((Nkall, No, No), dtype=complex)
    for i in range(Nkall):
        k, f = invk[i, 0] - 1, invk[i, 1]
        U = Uf[:, :, k]
        if f == 0:
            L = np.conj(U).T
            R = sl[:, None] * U[inv, :]
        elif f == 1:
            L = (-sl[None, :] * U[inv, :].T)
            R = np.conj(U)
        else:
            continue
        out[i] = L @ Df[i, 0] @ R
    return np.ascontiguousarray(out.T)


# --------------------------------------------------------------------------- #
#  nonlinear FLEX-Eliashberg (ffeliash.f90)
# --------------------------------------------------------------------------- #
def _full_ff_leg(Ff, pos_d, neg_d, pos_r, neg_r, c2, cj, kmap, invk, Nx, Ny, Nz):
    """One leg of the anomalous bubble on the full grid (ffeliash.f90:556-580):
    slot j <- F(k_irr, w_j, pos), slot M-1-j <- c2 F(k_irr, w_j, neg), the
    (a, b) pairs pos_d / neg_d at flag-0 points and pos_r / neg_r at -k points
    exactly as the reference indexes them; cj: conjugated leg (F^+14)."""
    Nk, Nw = Ff.shape[:2]
    M = 2 * Nw
    out = np.zeros((Nx * Ny * Nz, M), dtype=complex)
    irr = invk[:, 0] - 1
    fl = invk[:, 1]
    sp = _spatial(kmap, Nx, Ny, Nz)
    d = fl == 0
    r = fl == 1
    out[sp[d], :Nw] = Ff[irr[d], :, pos_d[0], pos_d[1]]
    out[sp[d], Nw:] = (c2 * Ff[irr[d], :, neg_d[0], neg_d[1]])[:, ::-1]
    out[sp[r], :Nw] = Ff[irr[r], :, pos_r[0], pos_r[1]]
    out[sp[r], Nw:] = (c2 * Ff[irr[r], :, neg_r[0], neg_r[1]])[:, ::-1]
    if cj:
        out = np.conj(out)
    return out.reshape(Nx, Ny, Nz, M)


def get_chi0_conv_ff(fk, kmap, invk, olist, prt, temp, Nx, Ny, Nz, singlet=True):
    """Anomalous bubble chi^FF(q) = -(T/Nkall) sum F^+14(k+q) F23(k) by FFT
    (ffeliash.f90:491-612): per irreducible pair (l, m) the legs F^+14 and F23
    with the parity / spin factors sgn = dprt = +1 (singlet) or -1 (triplet),
    the product with the (-r, -tau) reversal, the 1e-9 threshold at the
    gather and chi(chi_map(m,l)) = conj(chi(m,l)).  Python in/out orders."""
    Ff = np.asarray(fk).T
    prt = np.asarray(prt, dtype=float)
    ol = olist_f(olist)
    cmap, irr = get_chi_map(olist)
    Nk, Nw = Ff.shape[:2]
    Nchi = len(ol)
    Nkall = len(kmap)
    sgn = 1.0 if singlet else -1.0
    dprt = sgn
    weight = temp / Nkall
    chif = np.zeros((Nk, Nw, Nchi, Nchi), dtype=complex)
    for l, m in irr:
        l1, l2, m1, m2 = ol[l, 0], ol[l, 1], ol[m, 0], ol[m, 1]
        c14 = prt[l1] * prt[m2] * sgn * dprt
        c23 = prt[m1] * prt[l2] * sgn * dprt
        f14 = fft4(_full_ff_leg(Ff, (m2, l1), (l1, m2), (m1, l1), (l1, m2), c14, True, kmap, invk, Nx, Ny, Nz))
        f23 = fft4(_full_ff_leg(Ff, (l2, m1), (m1, l2), (l2, m2), (l2, m1), c23, False, kmap, invk, Nx, Ny, Nz))
        tmp = ifft4(-f14 * _reverse(f23))
        chif[:, :, m, l] = _threshold(_gather(tmp, kmap, invk, Nk, Nw) * weight)
        pm, pl = cmap[m, l]
        chif[:, :, pm, pl] = np.conj(chif[:, :, m, l])
    return np.ascontiguousarray(chif.T)


def get_chi0sc(Gk, fk, kmap, invk, olist, prt, temp, Nx, Ny, Nz, singlet=True):
    """chi0_S = chi^GG + chi^FF, chi0_C = chi^GG - chi^FF (ffeliash.f90:459-489)."""
    gg = get_chi0_conv(Gk, kmap, invk, olist, temp, Nx, Ny, Nz)
    ff = get_chi0_conv_ff(fk, kmap, invk, olist, prt, temp, Nx, Ny, Nz, singlet)
    return gg + ff, gg - ff


def mkV_flex_nosoc(chiS, chiC, Smat, Cmat, singlet=True):
    """V_sigma and V_Delta of the superconducting loop (ffeliash.f90:386-457):
    chi_s = (I - chi0_S S)^-1 chi0_S S, chi_c = (I + chi0_C C)^-1 chi0_C C,
    V_sigma = 3/2 S - 1/2 C + 3/2 S chi_s + 1/2 C chi_c
              - 1/4 (S + C)(chi0_S S + chi0_C C),
    V_Delta = (S + C)/2 + 3/2 S chi_s - 1/2 C chi_c (singlet) or
              (S - C)/2 - 1/2 S chi_s - 1/2 C chi_c (triplet).  Python orders."""
    AS = np.asarray(chiS).T
    AC = np.asarray(chiC).T
    S = np.asarray(Smat).T
    C = np.asarray(Cmat).T
    I = np.eye(S.shape[0])
    XS = AS @ S
    XC = AC @ C
    chis = np.linalg.solve(I - XS, XS)
    chic = np.linalg.solve(I + XC, XC)
    Vs = (1.5 * S - 0.5 * C) + 1.5 * (S @ chis) + 0.5 * (C @ chic) - 0.25 * ((S + C) @ (XS + XC))
    if singlet:
        Vd = 0.5 * (S + C) + 1.5 * (S @ chis) - 0.5 * (C @ chic)
    else:
        Vd = 0.5 * (S - C) - 0.5 * (S @ chis) - 0.5 * (C @ chic)
    return np.ascontiguousarray(Vs.T), np.ascontiguousarray(Vd.T)


def _hermitize_f(F):
    No = F.shape[-1]
    for l in range(No):
        F[..., l, l] = F[..., l, l].real
        for m in range(l + 1, No):
            F[..., l, m] = np.conj(F[..., m, l])
    return F


def mkgreliah_trs_nsoc(Gk0, fk, delta):
    """G = G0 (I + Delta F^dagger) per (k, iw) (ffeliash.f90:315-343)."""
    G0 = np.asarray(Gk0).T
    Ff = np.asarray(fk).T
    Df = np.asarray(delta).T
    No = G0.shape[-1]
    G = G0 @ (np.eye(No) + Df @ np.conj(np.swapaxes(Ff, -1, -2)))
    return np.ascontiguousarray(G.T)


def mkfkeliash_trs_nsoc(Gk0, Gk, delta):
    """F = -G0 Delta G^dagger, then diagonal real and upper = conj(lower)
    (ffeliash.f90:345-384)."""
    G0 = np.asarray(Gk0).T
    Gf = np.asarray(Gk).T
    Df = np.asarray(delta).T
    F = -(G0 @ (Df @ np.conj(np.swapaxes(Gf, -1, -2))))
    return np.ascontiguousarray(_hermitize_f(F).T)


def eliashberg(delta, Gk, hamk, Smat, Cmat, olist, prt, kmap, invk, mu, temp, eps, Nx, Ny, Nz, itemax, gap_sym,
               sw_sigma=False, sw_Vconst=False, m_diis=5, gap_min=1.0e-2, sw_amp_newton=True, info=None):
    """Non-linear FLEX-Eliashberg loop (ffeliash.f90:2-313), step for step:
    gap equation, convergence and gap-collapse tests, amplitude phase
    (probe / pitchfork fit / anchoring) or Pulay-DIIS (linear mixing pp = 0.3
    as fallback), FLEX Sigma and the dressed G0, the superconducting Dyson
    updates of G and F, chi0_S / chi0_C and the vertices.  Python in/out
    orders; returns (delta, sigmak, Gk) where Gk is the final normal G (the
    reference's Gk argument is intent(inout))."""
    D = np.array(np.asarray(delta).T, dtype=complex)
    G = np.array(np.asarray(Gk).T, dtype=complex)
    Nk, Nw, No, _ = D.shape
    Nkall = len(kmap)
    singlet = gap_sym >= 0
    weight = temp / Nkall
    pp, eps_reg, amp_tol, n_polish_max = 0.3, 1.0e-12, 2.0e-2, 10
    m_eff = max(int(m_diis), 1)
    xout = [None] * m_eff
    res = [None] * m_eff
    n_hist = i_hist = 0
    amp_phase, npts, n_polish, n_anchor = 1, 0, 0, 0
    a1 = a2 = g1 = g2 = 0.0
    a_cur = np.sqrt(np.vdot(D, D).real) + 1.0e-300
    v0 = D / a_cur
    if not sw_amp_newton:
        amp_phase = 2
    P = lambda X: np.ascontiguousarray(X.T)      # Fortran-order work array -> Python order
    Fo = lambda X: np.asarray(X).T               # Python order -> Fortran view
    G0d = G.copy()
    delta0_max = np.abs(D).max() + 1.0e-30
    if sw_Vconst:
        fk = np.zeros_like(G)
        cS, cC = get_chi0sc(P(G), P(fk), kmap, invk, olist, prt, temp, Nx, Ny, Nz, singlet)
        fk = Fo(mkfk_trs_nsoc(P(G), P(D))).copy()
    else:
        fk = Fo(mkfk_trs_nsoc(P(G), P(D))).copy()
        cS, cC = get_chi0sc(P(G), P(fk), kmap, invk, olist, prt, temp, Nx, Ny, Nz, singlet)
    stoner = ckchi(cS, Smat, Cmat)
    Vs, Vd = mkV_flex_nosoc(cS, cC, Smat, Cmat, singlet)
    sig = np.zeros_like(G)
    it_done = 0
    status = "itemax"
    for it in range(1, itemax + 1):
        it_done = it
        nd = mkdelta_nsoc(P(fk), Vd, Smat, Cmat, kmap, invk, prt, olist, Nx, Ny, Nz, singlet) * weight  # Fortran order
        diff = np.abs(nd - D).max()
        dnorm = np.abs(D).max() + 1.0e-30
        new_max = np.abs(nd).max()
        if info is not None:
            info.setdefault("trace", []).append((it, diff / dnorm, new_max, amp_phase))
        if diff / dnorm < eps:
            D = nd.copy()
            status = "converged"
            break
        if gap_min > 0.0 and new_max < gap_min * delta0_max:
            D = nd.copy()
            status = "collapsed"
            break
        if sw_amp_newton and amp_phase == 2 and n_polish >= n_polish_max:
            a_cur = np.sqrt(np.vdot(D, D).real) + 1.0e-300
            v0 = D / a_cur
            npts, a1, a2, g1, g2 = 0, 0.0, 0.0, 0.0, 0.0
            amp_phase = 1
            n_anchor += 1
        if amp_phase == 1:
            proj = np.vdot(v0, nd).real
            g_cur = proj / a_cur
            a2, g2, a1, g1 = a1, g1, a_cur, g_cur
            npts = min(npts + 1, 2)
            if npts >= 2 and abs(a1 * a1 - a2 * a2) > 1.0e-30:
                cfit = (g1 - g2) / (a2 * a2 - a1 * a1)
                lam0 = g1 + cfit * a1 * a1
                if cfit > 1.0e-12 and lam0 > 1.0:
                    a_nxt = np.sqrt((lam0 - 1.0) / cfit)
                    a_nxt = min(max(a_nxt, 0.3 * a_cur), 3.0 * a_cur)
                else:
                    a_nxt = 0.5 * a_cur
            else:
                a_nxt = 0.9 * a_cur if abs(g_cur - 1.0) < 0.2 else 0.6 * a_cur
            if abs(g_cur - 1.0) < amp_tol:
                amp_phase = 2
                a_nxt = a_cur
                n_polish = 0
                n_hist = i_hist = 0
            D = a_nxt * v0
            a_cur = a_nxt
        else:
            i_hist = i_hist % m_eff + 1
            n_hist = min(n_hist + 1, m_eff)
            n = n_hist
            xout[i_hist - 1] = nd.copy()
            res[i_hist - 1] = nd - D
            do_pulay = n >= 2 and m_eff >= 2
            if do_pulay:
                idx = [(i_hist - n + ih - 1) % m_eff for ih in range(1, n + 1)]
                B = np.zeros((n + 1, n + 1))
                for a in range(n):
                    for b in range(n):
                        B[a, b] = np.vdot(res[idx[a]], res[idx[b]]).real
                    B[a, a] += eps_reg
                    B[a, n] = B[n, a] = 1.0
                rhs = np.zeros(n + 1)
                rhs[n] = 1.0
                if info is not None:
                    info.setdefault("diis_cond", []).append(float(np.linalg.cond(B)))
                try:
                    c = np.linalg.solve(B, rhs)
                except np.linalg.LinAlgError:
                    do_pulay = False
            if do_pulay:
                D = np.zeros_like(D)
                for a in range(n):
                    D = D + c[a] * xout[idx[a]]
            else:
                D = pp * nd + (1.0 - pp) * D
            n_polish += 1
        if sw_sigma:
            sig = Fo(calc_sigma(P(G), Vs, Smat, Cmat, kmap, invk, olist, temp, Nx, Ny, Nz)).copy()
            G0d = Fo(gen_green(P(sig), hamk, mu, temp)).copy()
        G = Fo(mkgreliah_trs_nsoc(P(G0d), P(fk), P(D))).copy()
        fk = Fo(mkfkeliash_trs_nsoc(P(G0d), P(G), P(D))).copy()
        if not sw_Vconst:
            cS, cC = get_chi0sc(P(G), P(fk), kmap, invk, olist, prt, temp, Nx, Ny, Nz, singlet)
            stoner = ckchi(cS, Smat, Cmat)
            Vs, Vd = mkV_flex_nosoc(cS, cC, Smat, Cmat, singlet)
    if info is not None:
        info.update(iters=it_done, status=status, anchors=n_anchor, stoner=stoner)
    return P(D), P(sig), P(G)
