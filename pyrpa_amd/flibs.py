"""ctypes host layer of libgmod.so, mirroring the reference's libs/flibs wrappers.

Each function keeps the reference wrapper's name, argument order, defaults,
returned shapes and C-order array layouts (Python shapes are the reversed
Fortran shapes), so a caller of ``libs.flibs`` can switch by changing the
import.  Differences: errors raise ``RuntimeError`` instead of terminating the
process (the reference prints and STOPs), and the GPU library is required.
"""
from ctypes import POINTER, byref, c_bool, c_char_p, c_double, c_int, c_int32, c_int64, c_size_t, c_void_p

import numpy as np

from ._loader import check, dbl, i64, lib

_nd = np.ctypeslib.ndpointer
_z = _nd(dtype=np.complex128)
_d = _nd(dtype=np.float64)
_i = _nd(dtype=np.int64)
_i32 = _nd(dtype=np.int32)
_pi = POINTER(c_int64)
_pd = POINTER(c_double)
_pb = POINTER(c_bool)

_SIGS = {
    # libs/flibs/_selfenergy.py:7-37
    "get_vsigma_flex_nosoc_": [_z, _d, _d, _pi, _pi, _pi],
    "mkself": [_z, _pd, _d, _d, _i, _i, _i, _z, _d, _z, _pd, _pd, _pd, _pi, _pd, _pd, _pi, _pi, _pi, _pi,
               _pi, _pi, _pi, _pi, _pi, _pb, _pb, _pi, _pb, _pb, _pd],
    # libs/flibs/_susceptibility.py:7-204
    "get_chi0_conv_": [_z, _z, _i, _i, _i32, _i32, _i, _pd, _pi, _pi, _pi, _pi, _pi, _pi, _pi, _pi],
    "get_chi0": [_z, _d, _d, _z, _i, _i, _i, _pd, _pi, _pi, _pi, _pi, _pi, _pi, _pi, _pi, _pd],
    "get_chi0_tail": [_z, _d, _d, _z, _d, _z, _i, _i, _i, _pd, _pd, _pi, _pi, _pi, _pi, _pi, _pi, _pi, _pi,
                      _pd],
    "get_chis": [_z, _z, _d, _pi, _pi],
    "get_chis_chic": [_z, _z, _z, _d, _d, _pi, _pi, _pi],
    "get_eig_or_tr_chi": [_z, _z, _i, _pi, _pi, _pi, _pb],
    # addition: chi_s / chi_c at every (q, i nu)
    "gmod_get_chis_chic_all": [c_void_p, c_void_p, _z, _d, _d, _pi, _pi, _pi],
    # libs/flibs/_susceptibility.py:22-62 (real-frequency maps), _hamiltonian.py:79-85
    "get_chi_irr": [_z, _z, _d, _d, _i, _i, _d, _pi, _pi, _pi, _pi, _pd, _pd, _pd],
    "chiq_map": [_z, _z, _z, _d, _d, _d, _d, _i, _pd, _pd, _pd, _pd, _pi, _pi, _pi, _pi, _pi],
    "get_tr_chi": [_z, _z, _z, _z, _z, _i, _pi, _pi, _pi],
    "get_qshift_": [_d, _d, _i, _pi],
    # libs/flibs/_green.py:54-78 (Pade continuation)
    "get_a_": [_z, _z, _z, _pi],
    "get_qp_": [_z, _z, _z, _z, _z, _pi, _pi],
    "pade_with_trace": [_z, _z, _z, _z, _pi, _pi, _pi, _pi],
    "pade_analytic_continuation_arrays_": [_z, _z, _z, _z, _pi, _pi, _pi],
    # libs/flibs/_green.py:7-27
    "gen_green0_": [_z, _d, _z, _pd, _pd, _pi, _pi, _pi],
    "gen_green_inv_": [_z, _z, _z, _pd, _pd, _pi, _pi, _pi],
    "getinv_": [_z, _pi, _pi, _pi],
    # libs/flibs/_eliashberg.py:7-143
    "lin_eliash": [_z, _z, _z, _z, _d, _d, _d, _i, _d, _i, _i, _pd, _pd, _pi, _pi, _pi, _pi, _pi, _pi, _pi,
                   _pi, _pi, _pi, _pi, _pd],
    "mkfk_trs_nsoc_": [_z, _z, _z, _pi, _pi, _pi],
    "get_initial_delta_": [_z, _d, _z, _i, _i, _pi, _pi, _pi, _pi, _pi],
    "conv_delta_orb_to_band": [_z, _z, _z, _d, _i, _pi, _pi, _pi, _pi, _pi],
    "remap_delta_": [_z, _z, _d, _i, _pi, _pi, _pi, _pi, _pi],
    # libs/flibs/_eliashberg.py:230-304 (nonlinear FLEX-Eliashberg) and the
    # loop's internal bubbles / vertices (ffeliash.f90:386-489)
    "eliashberg": [_z, _z, _z, _z, _d, _d, _i, _d, _i, _i, _pd, _pd, _pd, _pi, _pi, _pi, _pi, _pi, _pi, _pi,
                   _pi, _pi, _pi, _pb, _pb, _pi, _pd, _pb],
    "gmod_eliashberg_stats": [_d, c_int],
    "get_chi0sc_": [_z, _z, _z, _z, _i, _i, _i32, _i32, _i, _d, _pd, _pi, _pi, _pi, _pi, _pi, _pi, _pi, _pi, _pb],
    "mkv_flex_nosoc_": [_z, _z, _d, _d, _pi, _pi, _pi, _pb],
    # libs/flibs/_interaction.py:7-45
    "get_scmat": [_d, _d, _i, _i, _pd, _pd, _pi],
    "get_scmat_orb": [_d, _d, _i, _i, _d, _d, _pi, _pi],
    "get_vmat_soc": [_d, _i, _i, _i, _i, _pd, _pd, _pi, _pi],
    "get_vmat_soc_orb": [_d, _i, _i, _i, _i, _d, _d, _pi, _pi],
    # spinful basis: libs/flibs/_selfenergy.py:38-60, _susceptibility.py:141-173
    "mkself_soc": [_z, _pd, _d, _i, _i, _i, _i, _i, _z, _d, _z, _pd, _pd, _pd, _pi, _pd, _pd, _pi, _pi, _pi, _pi,
                   _pi, _pi, _pi, _pi, _pi, _pb, _pb, _pi, _pb, _pd],
    "get_chi0_soc": [_z, _d, _d, _i, _d, _z, _i, _i, _i, _i, _i, _pd, _pi, _pi, _pi, _pi, _pi, _pi, _pi, _pi],
    "get_chis_chic_soc": [_z, _z, _z, _z, _d, _i, _i, _i, _i, _pi, _pi, _pi, _pi],
    "gmod_v_soc_flex": [_z, _d, _pi, _pi, _pi],
    # libs/flibs/_eliashberg.py:29-52, 96-115
    "lin_eliash_soc": [_z, _z, _z, _z, _d, _d, _d, _d, _d, _i, _i, _i, _i, _i, _i, _pd, _pd, _pi, _pi, _pi, _pi,
                       _pi, _pi, _pi, _pi, _pi, _pi, _pi],
    "mkfk_trs_soc_": [_z, _z, _z, _d, _i, _i, _i, _pi, _pi, _pi, _pi, _pi],
    "conv_delta_orb_to_band_soc": [_z, _z, _z, _i, _i, _i, _pi, _pi, _pi, _pi],
    # libs/flibs/_hamiltonian.py:7-106
    "openmp_params": [_pi, _pb],
    "gen_ham": [_z, _d, _z, _d, _pi, _pi, _pi],
    "get_eig": [_d, _z, _z, _pi, _pi],
    "get_ffermi": [_d, _d, _pd, _pd, _pi, _pi],
    "generate_irr_kpoint_inv": [_d, _i, _i, _pi, _pi, _pi, _pi],
    "get_kweight": [_d, _i, _pi, _pi],
}

_bound = False


def _L():
    global _bound
    L = lib()
    if not _bound:
        for name, args in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = None
        L.gmod_last_error.argtypes = [c_char_p, c_int]
        L.gmod_last_error.restype = c_int
        L.gmod_set_precision.argtypes = [c_int]
        L.gmod_get_precision.restype = c_int
        L.gmod_set_sigma_seed.argtypes = [_z, c_double, c_double, c_int64]
        L.gmod_flex_stats.argtypes = [_d, c_int]
        L.gmod_flex_stats.restype = c_int
        L.gmod_regrid_sigma.argtypes = [_z, _z, _i, _i, c_int64, c_int64, c_int64, c_int64, c_int64, _i, _i, c_int64,
                                        c_int64, c_int64, c_int64, c_int64, c_int64, c_double]
        L.gmod_regrid_sigma.restype = None
        L.gmod_wannier_r.argtypes = [_z, _z, _i, _i] + [c_int64] * 7 + [c_int]
        L.gmod_wannier_r.restype = None
        L.gmod_last_lambda.argtypes = []
        L.gmod_last_lambda.restype = c_double
        L.gmod_profile.argtypes = [c_int]
        L.gmod_profile_read.argtypes = [_d, c_int]
        L.gmod_profile_read.restype = c_int
        L.gmod_set_verbose.argtypes = [c_int]
        L.gmod_pool_bytes.restype = c_double
        L.gmod_fft_device.argtypes = [c_void_p, c_void_p] + [c_int] * 8
        L.gmod_device_alloc.argtypes = [c_size_t]
        L.gmod_device_alloc.restype = c_void_p
        L.gmod_device_free.argtypes = [c_void_p]
        L.gmod_device_free.restype = None
        L.gmod_memcpy.argtypes = [c_void_p, c_void_p, c_size_t, c_int]
        L.gmod_memcpy.restype = c_int
        L.gmod_eig_or_tr_chi_noalloc.argtypes = [c_void_p, c_void_p, c_void_p, c_int64, c_int64, c_int64, c_int]
        L.gmod_eig_or_tr_chi_noalloc.restype = c_int
        L.gmod_fft_device.restype = c_int
        _bound = True
    return L


def _call(name, *args):
    getattr(_L(), name)(*args)
    check()


# =========================================================================== hamiltonian / k-mesh
def omp_params():
    """(device count, True if a HIP device is present)  [_hamiltonian.py:122-132]"""
    n = c_int64()
    ok = c_bool()
    _call("openmp_params", byref(n), byref(ok))
    return n.value, ok.value


def gen_ham(klist, ham_r, rvec, Ovl_r=None):
    """H(k) = sum_R H(R) exp(-2 pi i k.R)  [_hamiltonian.py:134-158]"""
    if Ovl_r is not None:
        raise NotImplementedError("overlap matrices (MLO bases) are outside the FLEX/Eliashberg path")
    klist = np.ascontiguousarray(klist, dtype=np.float64)
    rvec = np.ascontiguousarray(rvec, dtype=np.float64)
    ham_r = np.ascontiguousarray(ham_r, dtype=np.complex128)
    Nk, Nr = len(klist), len(rvec)
    Norb = ham_r.shape[1]
    hamk = np.zeros((Nk, Norb, Norb), dtype=np.complex128)
    _call("gen_ham", hamk, klist, ham_r, rvec, i64(Nk), i64(Nr), i64(Norb))
    return hamk


def get_eig(hamk, Ovlk=None, sw=True):
    """eig[Nk,No], uni[Nk,No,No]  [_hamiltonian.py:160-184]"""
    if Ovlk is not None:
        raise NotImplementedError("generalised eigenproblem is outside the FLEX/Eliashberg path")
    hamk = np.ascontiguousarray(hamk, dtype=np.complex128)
    Nk, Norb = len(hamk), hamk.shape[1]
    eig = np.zeros((Nk, Norb), dtype=np.float64)
    uni = np.zeros((Nk, Norb, Norb), dtype=np.complex128)
    _call("get_eig", eig, uni, hamk, i64(Nk), i64(Norb))
    return (eig, uni) if sw else eig


def get_ffermi(eig, mu, temp):
    """Fermi occupations  [_hamiltonian.py:186-200]"""
    eig = np.ascontiguousarray(eig, dtype=np.float64)
    Nk, Norb = eig.shape
    f = np.zeros((Nk, Norb), dtype=np.float64)
    _call("get_ffermi", f, eig, dbl(mu), dbl(temp), i64(Nk), i64(Norb))
    return f


def gen_irr_k_TRS(Nx, Ny, Nz):
    """klist[Nk,3], kmap[Nkall,3], invk[Nkall,3]  [_hamiltonian.py:356-382]"""
    Nkall = Nx * Ny * Nz
    if Nkall % 2 == 0:
        if Nz % 2 == 0:
            Nk = int(Nkall / 2 + 4) if ((Nx % 2) == 0 and (Ny % 2) == 0) else (
                int(Nkall / 2 + 2) if ((Nx % 2) == 0 or (Ny % 2) == 0) else int(Nkall / 2 + 1))
        else:
            Nk = int(Nkall / 2 + 2) if ((Nx % 2) == 0 and (Ny % 2) == 0) else int(Nkall / 2 + 1)
    else:
        Nk = int((Nkall + 1) / 2)
    klist = np.zeros((Nk, 3), dtype=np.float64)
    kmap = np.zeros((Nkall, 3), dtype=np.int64)
    invk = np.zeros((Nkall, 3), dtype=np.int64)
    _call("generate_irr_kpoint_inv", klist, kmap, invk, i64(Nk), i64(Nx), i64(Ny), i64(Nz))
    return klist, kmap, invk


def gen_kpoint_weight(invk, Nk):
    """[_hamiltonian.py:384-395]"""
    w = np.zeros(Nk, dtype=np.float64)
    _call("get_kweight", w, np.ascontiguousarray(invk, dtype=np.int64), i64(Nk), i64(len(invk)))
    return w


# =========================================================================== interaction
def gen_SCmatrix(olist, site, U, J):
    """[_interaction.py:47-62]"""
    Nchi = len(olist)
    S = np.zeros((Nchi, Nchi), dtype=np.float64)
    C = np.zeros((Nchi, Nchi), dtype=np.float64)
    _call("get_scmat", S, C, olist, np.asarray(site, dtype=np.int64), dbl(U), dbl(J), i64(Nchi))
    return S, C


def gen_SCmatrix_orb(olist, site, Umat, Jmat):
    """[_interaction.py:64-81]"""
    Nchi, Norb = len(olist), len(Umat)
    S = np.zeros((Nchi, Nchi), dtype=np.float64)
    C = np.zeros((Nchi, Nchi), dtype=np.float64)
    _call("get_scmat_orb", S, C, olist, np.asarray(site, dtype=np.int64), np.asarray(Umat, dtype=np.float64),
          np.asarray(Jmat, dtype=np.float64), i64(Nchi), i64(Norb))
    return S, C


# =========================================================================== Green functions
def gen_Green0(eig, uni, mu, temp, Nw):
    """G0[No,No,Nw,Nk]  [_green.py:80-97]"""
    Nk, Norb = np.shape(eig)
    Gk = np.zeros((Norb, Norb, Nw, Nk), dtype=np.complex128)
    _call("gen_green0_", Gk, np.ascontiguousarray(eig), np.ascontiguousarray(uni), dbl(mu), dbl(temp), i64(Nk),
          i64(Nw), i64(Norb))
    return Gk


def gen_green(selfen, hamk, mu, temp):
    """G = [i w + mu - H - Sigma]^-1  [_green.py:99-117]"""
    Nk, Norb = len(hamk), hamk.shape[1]
    Nw = selfen.shape[2]
    Gk = np.zeros((Norb, Norb, Nw, Nk), dtype=np.complex128)
    _call("gen_green_inv_", Gk, np.ascontiguousarray(selfen), np.ascontiguousarray(hamk), dbl(mu), dbl(temp),
          i64(Nk), i64(Nw), i64(Norb))
    _call("getinv_", Gk, i64(Nk), i64(Nw), i64(Norb))
    return Gk


# =========================================================================== susceptibility
def _get_chi_map(olist):
    """chi_map (2,Nchi,Nchi) and irr_chi (2,P) int32 with the memory layout of
    the Fortran chi_map(Nchi,Nchi,2), irr_chi(P,2)  [_susceptibility.py:206-242,
    algorithm of fchi.f90:421-481]"""
    ol = np.asarray(olist, dtype=np.int64)
    Nchi = len(ol)
    cmap = np.zeros((2, Nchi, Nchi), dtype=np.int32)
    irr_flag = np.zeros((Nchi, Nchi), dtype=np.int32)
    for l1 in range(Nchi):
        for m1 in range(Nchi):
            key = (ol[l1, 0], ol[l1, 1], ol[m1, 0], ol[m1, 1])
            hit = None
            for l2 in range(Nchi):
                for m2 in range(Nchi):
                    if key == (ol[m2, 1], ol[m2, 0], ol[l2, 1], ol[l2, 0]):
                        hit = (l2, m2)
                        break
                if hit:
                    break
            if hit:
                l2, m2 = hit
                cmap[0, l1, m1] = m2 + 1
                cmap[1, l1, m1] = l2 + 1
                if irr_flag[l2, m2] == 0:
                    irr_flag[l1, m1] = 1
    irr = np.zeros((2, Nchi * (Nchi + 1) // 2), dtype=np.int32)
    it = 0
    for l1 in range(Nchi):
        for m1 in range(Nchi):
            if irr_flag[l1, m1] == 1 and it < irr.shape[1]:
                irr[0, it], irr[1, it] = l1 + 1, m1 + 1
                it += 1
    return cmap, irr


def get_chi0_conv(Gk, kmap, invk, olist, temp, Nx, Ny, Nz):
    """chi0 by FFT convolution, sharp cutoff  [_susceptibility.py:244-269]"""
    Nkall, Nk = len(kmap), Gk.shape[3]
    Norb, Nchi, Nw = len(Gk), len(olist), Gk.shape[2]
    chi = np.zeros((Nchi, Nchi, Nw, Nk), dtype=np.complex128)
    chi_map, irr_chi = _get_chi_map(olist)
    _call("get_chi0_conv_", chi, np.ascontiguousarray(Gk), kmap, invk, irr_chi, chi_map, olist, dbl(temp),
          i64(Nx), i64(Ny), i64(Nz), i64(Nw), i64(Nk), i64(Nkall), i64(Norb), i64(Nchi))
    return chi


def get_chi0(Smat, Cmat, Gk, olist, kmap, invk, temp, Nx, Ny, Nz):
    """(chi0[Nchi,Nchi,Nw,Nk], Stoner factor)  [_susceptibility.py:432-458]"""
    Norb, Nchi = len(Gk), len(olist)
    Nkall, Nk, Nw = len(kmap), Gk.shape[3], Gk.shape[2]
    chi = np.zeros((Nchi, Nchi, Nw, Nk), dtype=np.complex128)
    st = c_double(0.0)
    _call("get_chi0", chi, Smat, Cmat, np.ascontiguousarray(Gk), kmap, invk, olist, dbl(temp), i64(Nx), i64(Ny),
          i64(Nz), i64(Nw), i64(Nk), i64(Nkall), i64(Norb), i64(Nchi), byref(st))
    return chi, st.value


def get_chi0_tail(Smat, Cmat, Gk, eig, uni, olist, kmap, invk, mu, temp, Nx, Ny, Nz):
    """tail-corrected chi0  [_susceptibility.py:460-494]"""
    Norb, Nchi = len(Gk), len(olist)
    Nkall, Nk, Nw = len(kmap), Gk.shape[3], Gk.shape[2]
    chi = np.zeros((Nchi, Nchi, Nw, Nk), dtype=np.complex128)
    st = c_double(0.0)
    _call("get_chi0_tail", chi, Smat, Cmat, np.ascontiguousarray(Gk), np.ascontiguousarray(eig),
          np.ascontiguousarray(uni), kmap, invk, olist, dbl(mu), dbl(temp), i64(Nx), i64(Ny), i64(Nz), i64(Nw),
          i64(Nk), i64(Nkall), i64(Norb), i64(Nchi), byref(st))
    return chi, st.value


def get_chis(chi0, Smat):
    """[_susceptibility.py:419-430]"""
    Nchi, Nw = len(Smat), len(chi0)
    chis = np.zeros((Nw, Nchi, Nchi), dtype=np.complex128)
    _call("get_chis", chis, np.ascontiguousarray(chi0), Smat, i64(Nchi), i64(Nw))
    return chis


def get_chis_chic(chi, Smat, Cmat):
    """static RPA chi_s, chi_c [Nchi,Nchi,Nk]  [_susceptibility.py:568-583]"""
    Nk, Nw, Nchi = chi.shape[3], chi.shape[2], len(Smat)
    chis = np.zeros((Nchi, Nchi, Nk), dtype=np.complex128)
    chic = np.zeros((Nchi, Nchi, Nk), dtype=np.complex128)
    _call("get_chis_chic", chis, chic, np.ascontiguousarray(chi), Smat, Cmat, i64(Nk), i64(Nw), i64(Nchi))
    return chis, chic


def get_chis_chic_all(chi, Smat, Cmat, spin=True, charge=True):
    """RPA chi_s and chi_c at every (q, i nu), [Nchi,Nchi,Nw,Nk] each like chi
    (None for a susceptibility not asked for): the algebra of get_chis /
    get_chis_chic applied to the whole batch (addition; no single reference
    wrapper)"""
    Nk, Nw, Nchi = chi.shape[3], chi.shape[2], len(Smat)
    chis = np.zeros((Nchi, Nchi, Nw, Nk), dtype=np.complex128) if spin else None
    chic = np.zeros((Nchi, Nchi, Nw, Nk), dtype=np.complex128) if charge else None
    _call("gmod_get_chis_chic_all", None if chis is None else chis.ctypes.data,
          None if chic is None else chic.ctypes.data, np.ascontiguousarray(chi), np.ascontiguousarray(Smat),
          np.ascontiguousarray(Cmat), i64(Nk), i64(Nw), i64(Nchi))
    return chis, chic


def get_qshift(klist, qpoint):
    """1-based index of k + q for every k of klist[Nk,3]  [_hamiltonian.py:330-341]"""
    kl = np.ascontiguousarray(klist, dtype=np.float64)
    Nk = len(kl)
    qshift = np.zeros(Nk, dtype=np.int64)
    _call("get_qshift_", np.ascontiguousarray(qpoint, dtype=np.float64), kl, qshift, i64(Nk))
    return qshift


def get_chi_irr(uni, eig, ffermi, qshift, olist, wlist, idelta, temp):
    """chi0(q, w + i idelta) [Nw, Nchi, Nchi] by the band sum  [_susceptibility.py:271-293]"""
    Nk, Nw = len(eig), len(wlist)
    Norb, Nchi = eig.shape[1], len(olist)
    chi = np.zeros((Nw, Nchi, Nchi), dtype=np.complex128)
    eps = idelta * 1e-3
    _call("get_chi_irr", chi, np.ascontiguousarray(uni), np.ascontiguousarray(eig), np.ascontiguousarray(ffermi),
          np.ascontiguousarray(qshift, dtype=np.int64), olist, np.ascontiguousarray(wlist, dtype=np.float64),
          i64(Nchi), i64(Norb), i64(Nk), i64(Nw), dbl(idelta), dbl(eps), dbl(temp))
    return chi


def chis_qmap(uni, eig, ffermi, klist, Smat, olist, Nx, Ny, temp, ecut, idelta):
    """(chis, chi) [Nx, Ny]: traces over the orbital-diagonal pairs of chi_s
    and chi0 at w = ecut on the q plane  [_susceptibility.py:295-324]"""
    Nk = len(eig)
    Norb, Nchi = eig.shape[1], len(olist)
    chi = np.zeros((Nx, Ny), dtype=np.complex128)
    chis = np.zeros((Nx, Ny), dtype=np.complex128)
    eps = idelta * 1e-3
    _call("chiq_map", chis, chi, np.ascontiguousarray(uni), np.ascontiguousarray(eig), np.ascontiguousarray(ffermi),
          np.ascontiguousarray(klist, dtype=np.float64), np.ascontiguousarray(Smat), olist, dbl(temp), dbl(ecut),
          dbl(idelta), dbl(eps), i64(Nx), i64(Ny), i64(Nk), i64(Norb), i64(Nchi))
    return chis, chi


def get_tr_chi(chis, chi0, olist):
    """(trchis[Nw], trchi0[Nw], chis_orb[Nw, Norb+2])  [_susceptibility.py:326-344]"""
    Nchi, Nw, Norb = len(olist), len(chi0), int(np.asarray(olist).max())
    trchis = np.zeros(Nw, dtype=np.complex128)
    trchi0 = np.zeros(Nw, dtype=np.complex128)
    chis_orb = np.zeros((Nw, Norb + 2), dtype=np.complex128)
    _call("get_tr_chi", trchis, trchi0, chis_orb, np.ascontiguousarray(chis), np.ascontiguousarray(chi0), olist,
          i64(Nw), i64(Nchi), i64(Norb))
    return trchis, trchi0, chis_orb


def get_a(inp_data, xlist):
    """Pade continued-fraction coefficients [Np]  [_green.py:177-188]"""
    Np = len(inp_data)
    a = np.zeros(Np, dtype=np.complex128)
    _call("get_a_", a, np.ascontiguousarray(xlist, dtype=np.complex128),
          np.ascontiguousarray(inp_data, dtype=np.complex128), i64(Np))
    return a


def get_QP(a, xlist, wlist):
    """(Q, P) of the continued fraction at wlist, in the reference wrapper's
    return order  [_green.py:190-204]"""
    Nw, Np = len(wlist), len(a)
    Q = np.zeros(Nw, dtype=np.complex128)
    P = np.zeros(Nw, dtype=np.complex128)
    _call("get_qp_", P, Q, np.ascontiguousarray(a, dtype=np.complex128), np.ascontiguousarray(xlist, dtype=np.complex128),
          np.ascontiguousarray(wlist, dtype=np.complex128), i64(Nw), i64(Np))
    return Q, P


def pade_with_trace(A, iwlist, wlist):
    """[Nk, Nw]: Pade continuation of the orbital-diagonal blocks of
    A[Norb,Norb,Niw,Nk], summed  [_green.py:206-220]"""
    Nk, Nw, Niw = len(A.T), len(wlist), len(iwlist)
    Norb = A.shape[0]
    B = np.zeros((Nk, Nw), dtype=np.complex128)
    _call("pade_with_trace", np.ascontiguousarray(A, dtype=np.complex128), B,
          np.ascontiguousarray(iwlist, dtype=np.complex128), np.ascontiguousarray(wlist, dtype=np.complex128),
          i64(Nk), i64(Niw), i64(Nw), i64(Norb))
    return B


def pade_arrays(arrayin, iwlist, wlist):
    """[Nw, Nk] view of the Fortran arrayout(Nk,Nw): the continuation of every
    row of arrayin (Python [Niw, Nk], Fortran arrayin(Nk,Niw))
    (pade_analytic_continuation_arrays_, fpade.f90:87-106; no reference wrapper)"""
    Niw, Nk = np.shape(arrayin)
    Nw = len(wlist)
    out = np.zeros((Nw, Nk), dtype=np.complex128)
    _call("pade_analytic_continuation_arrays_", np.ascontiguousarray(arrayin, dtype=np.complex128), out,
          np.ascontiguousarray(iwlist, dtype=np.complex128), np.ascontiguousarray(wlist, dtype=np.complex128),
          i64(Nk), i64(Niw), i64(Nw))
    return out


def get_eig_or_tr_chi(chi, invk, sw_eig):
    """[_susceptibility.py:615-629]"""
    Nkall, Nk = len(invk), chi.shape[2]
    Nchi = chi.shape[0]
    chiq = np.zeros(Nkall, dtype=np.complex128)
    _call("get_eig_or_tr_chi", chiq, np.ascontiguousarray(chi), invk, i64(Nkall), i64(Nk), i64(Nchi),
          byref(c_bool(sw_eig)))
    return chiq


# =========================================================================== self-energy
def get_Vsigma_nosoc_flex(chi, Smat, Cmat):
    """V_sigma (chi modified in place, copy returned)  [_selfenergy.py:62-74]"""
    Nk, Nw, Nchi = chi.shape[3], chi.shape[2], len(Smat)
    _call("get_vsigma_flex_nosoc_", chi, Smat, Cmat, i64(Nk), i64(Nw), i64(Nchi))
    return chi.copy()


def mkself(Smat, Cmat, kmap, invk, olist, hamk, eig, uni, mu, fill, temp, Nw, Nx, Ny, Nz, sw_out, sw_in,
           sub_sigma=1, scf_loop=300, eps=1.0e-4, pp=0.3, m_diis=5, sw_rescale=False, sw_tail=False,
           sigma_scale=1.0):
    """FLEX self-consistency loop  [_selfenergy.py:76-127]"""
    Nkall, Nk, Nchi = len(kmap), len(hamk), len(Smat)
    Norb = hamk.shape[1]
    mu_self = c_double()
    sigmak = np.zeros((Norb, Norb, Nw, Nk), dtype=np.complex128)
    _call("mkself", sigmak, byref(mu_self), Smat, Cmat, kmap, invk, olist, np.ascontiguousarray(hamk),
          np.ascontiguousarray(eig), np.ascontiguousarray(uni), dbl(mu), dbl(fill), dbl(temp), i64(scf_loop),
          dbl(pp), dbl(eps), i64(Nkall), i64(Nk), i64(Nw), i64(Norb), i64(Nchi), i64(Nx), i64(Ny), i64(Nz),
          i64(sub_sigma), byref(c_bool(sw_out)), byref(c_bool(sw_in)), i64(m_diis), byref(c_bool(sw_rescale)),
          byref(c_bool(sw_tail)), dbl(sigma_scale))
    return sigmak, mu_self.value


# =========================================================================== spinful basis (SOC)
def gen_Vmatrix(olist, slist, site, invs, U, J):
    """SOC interaction vertex [Nchi, Nchi]  [_interaction.py:83-100]"""
    Nchi, Norb = len(olist), len(slist)
    V = np.zeros((Nchi, Nchi), dtype=np.float64)
    _call("get_vmat_soc", V, olist, np.asarray(slist, dtype=np.int64), np.asarray(site, dtype=np.int64),
          np.asarray(invs, dtype=np.int64), dbl(U), dbl(J), i64(Nchi), i64(Norb))
    return V


def gen_Vmatrix_orb(olist, slist, site, invs, Umat, Jmat):
    """[_interaction.py:102-119]"""
    Nchi, Norb = len(olist), len(Umat)
    V = np.zeros((Nchi, Nchi), dtype=np.float64)
    _call("get_vmat_soc_orb", V, olist, np.asarray(slist, dtype=np.int64), np.asarray(site, dtype=np.int64),
          np.asarray(invs, dtype=np.int64), np.asarray(Umat, dtype=np.float64), np.asarray(Jmat, dtype=np.float64),
          i64(Nchi), i64(Norb))
    return V


def get_chi0_soc(Vmat, Gk, olist, slist, kmap, invk, invs, temp, Nx, Ny, Nz):
    """(chi[Nchi,Nchi,Nw,Nk], sgnsig[Norb,Norb], sgnsig2[Nchi,Nchi], invschi[Nchi])
    [_susceptibility.py:496-528]"""
    Norb, Nchi = len(slist), len(olist)
    Nkall, Nk, Nw = len(kmap), Gk.shape[3], Gk.shape[2]
    chi = np.zeros((Nchi, Nchi, Nw, Nk), dtype=np.complex128)
    sgnsig = np.zeros((Norb, Norb), dtype=np.float64)
    sgnsig2 = np.zeros((Nchi, Nchi), dtype=np.float64)
    invschi = np.zeros(Nchi, dtype=np.int64)
    _call("get_chi0_soc", chi, sgnsig, sgnsig2, invschi, np.asarray(Vmat, dtype=np.float64), np.ascontiguousarray(Gk),
          kmap, invk, np.asarray(invs, dtype=np.int64), olist, np.asarray(slist, dtype=np.int64), dbl(temp), i64(Nx),
          i64(Ny), i64(Nz), i64(Nw), i64(Nk), i64(Nkall), i64(Nchi), i64(Norb))
    return chi, sgnsig, sgnsig2, invschi


def _soc_orb_list(olist, slist, invs):
    """1-based index of each chi pair among the spin-up orbital pairs
    (get_orb_list of _susceptibility.py:545-556): pair n = i nu + j of the
    spin-up orbitals up[] is (up[j], up[i]); a spin-down orbital counts as its
    spin-flipped partner."""
    ol = np.asarray(olist)
    sl = np.asarray(slist)
    inv = np.asarray(invs)
    up = np.nonzero(sl == 1)[0] + 1
    pos = {(int(up[j]), int(up[i])): i * len(up) + j + 1 for i in range(len(up)) for j in range(len(up))}
    out = np.zeros(len(ol), dtype=np.int64)
    for i, (a, b) in enumerate(ol):
        a = a if sl[a - 1] == 1 else inv[a - 1]
        b = b if sl[b - 1] == 1 else inv[b - 1]
        out[i] = pos.get((int(a), int(b)), 0)
    return out


def get_chis_chic_soc(chi, Vmat, olist, slist, invs):
    """(chic, chiszz, chispm) [Nchi/4, Nchi/4, Nk]  [_susceptibility.py:530-566]"""
    Norb, Nchi = len(slist), len(olist)
    Nk, Nw = chi.shape[3], chi.shape[2]
    n4 = Nchi // 4
    chic = np.zeros((n4, n4, Nk), dtype=np.complex128)
    chiszz = np.zeros((n4, n4, Nk), dtype=np.complex128)
    chispm = np.zeros((n4, n4, Nk), dtype=np.complex128)
    _call("get_chis_chic_soc", chic, chiszz, chispm, np.ascontiguousarray(chi), np.asarray(Vmat, dtype=np.float64),
          _soc_orb_list(olist, slist, invs), olist, np.asarray(slist, dtype=np.int64), np.asarray(invs, dtype=np.int64),
          i64(Nk), i64(Nw), i64(Nchi), i64(Norb))
    return chic, chiszz, chispm


def get_V_soc_flex(chi, Vmat):
    """-V + V (I + chi0 V)^-1 chi0 V per (q, i nu) (fUJ.f90:367-399; internal to
    mkself_soc in the reference): a new array [Nchi,Nchi,Nw,Nk]"""
    out = np.array(chi, dtype=np.complex128, order='C')
    Nchi, _, Nw, Nk = out.shape
    _call("gmod_v_soc_flex", out, np.asarray(Vmat, dtype=np.float64), i64(Nk), i64(Nw), i64(Nchi))
    return out


def linearized_eliashberg_soc(chi, Gk, uni, init_delta, Vmat, sgnsig, sgnsig2, plist, slist, olist, kmap, invk,
                              invs, invschi, Nx, Ny, Nz, temp, gap_sym, eps=1.0e-4, itemax=300, arnoldi_m=10):
    """delta[No,No,Nw,Nkall] on the full zone  [_eliashberg.py:183-223]; chi is
    overwritten with V_Delta; the eigenvalue (printed by the reference) is
    last_lambda()"""
    Norb, Nchi = len(slist), len(Vmat)
    Nkall, Nk, Nw = len(kmap), Gk.shape[3], Gk.shape[2]
    delta = np.zeros((Norb, Norb, Nw, Nkall), dtype=np.complex128)
    _call("lin_eliash_soc", delta, chi, np.ascontiguousarray(Gk), np.ascontiguousarray(uni),
          np.ascontiguousarray(init_delta, dtype=np.float64), np.asarray(Vmat, dtype=np.float64),
          np.asarray(sgnsig, dtype=np.float64), np.asarray(sgnsig2, dtype=np.float64),
          np.asarray(plist, dtype=np.float64), olist, np.asarray(slist, dtype=np.int64), kmap, invk,
          np.asarray(invs, dtype=np.int64), np.asarray(invschi, dtype=np.int64), dbl(temp), dbl(eps), i64(Nkall),
          i64(Nk), i64(Nw), i64(Nchi), i64(Norb), i64(Nx), i64(Ny), i64(Nz), i64(itemax), i64(gap_sym),
          i64(arnoldi_m))
    return delta


def last_lambda():
    """Eigenvalue of the last linearized_eliashberg_soc call (gmod_last_lambda)"""
    return float(_L().gmod_last_lambda())


def gen_Fk_soc(Gk, delta, invk, invs, slist, gap_sym):
    """F[No,No,Nw,Nkall]  [_eliashberg.py:357-376]"""
    Nkall, Nk, Nw, Norb = len(invk), Gk.shape[3], delta.shape[2], len(delta)
    Fk = np.zeros((Norb, Norb, Nw, Nkall), dtype=np.complex128)
    sl = np.asarray(slist, dtype=np.int64)
    sgnsig = np.outer(sl, sl).astype(np.float64)
    _call("mkfk_trs_soc_", Fk, np.ascontiguousarray(Gk), np.ascontiguousarray(delta), sgnsig, sl, invk,
          np.asarray(invs, dtype=np.int64), i64(Nkall), i64(Nk), i64(Nw), i64(Norb), i64(gap_sym))
    return Fk


def conv_delta_orb_to_band_soc(delta, uni, invk, invs, slist):
    """deltab[No,No,Nkall]  [_eliashberg.py:325-343]"""
    Nkall, Nk, Nw, Norb = len(invk), len(uni), delta.shape[2], len(delta)
    deltab = np.zeros((Norb, Norb, Nkall), dtype=np.complex128)
    _call("conv_delta_orb_to_band_soc", deltab, np.ascontiguousarray(delta), np.ascontiguousarray(uni), invk,
          np.asarray(invs, dtype=np.int64), np.asarray(slist, dtype=np.int64), i64(Norb), i64(Nkall), i64(Nk),
          i64(Nw))
    return deltab


def mkself_soc(Vmat, kmap, invk, invs, olist, slist, hamk, eig, uni, mu, fill, temp, Nw, Nx, Ny, Nz, sw_out, sw_in,
               sub_sigma=1, scf_loop=300, eps=1.0e-4, pp=0.3, m_diis=5, sw_rescale=False, sigma_scale=1.0):
    """FLEX loop in a spinful basis  [_selfenergy.py:129-174]"""
    Nkall, Nk, Nchi = len(kmap), len(hamk), len(Vmat)
    Norb = hamk.shape[1]
    mu_self = c_double()
    sigmak = np.zeros((Norb, Norb, Nw, Nk), dtype=np.complex128)
    _call("mkself_soc", sigmak, byref(mu_self), np.asarray(Vmat, dtype=np.float64), kmap, invk,
          np.asarray(invs, dtype=np.int64), olist, np.asarray(slist, dtype=np.int64), np.ascontiguousarray(hamk),
          np.ascontiguousarray(eig), np.ascontiguousarray(uni), dbl(mu), dbl(fill), dbl(temp), i64(scf_loop), dbl(pp),
          dbl(eps), i64(Nkall), i64(Nk), i64(Nw), i64(Norb), i64(Nchi), i64(Nx), i64(Ny), i64(Nz), i64(sub_sigma),
          byref(c_bool(sw_out)), byref(c_bool(sw_in)), i64(m_diis), byref(c_bool(sw_rescale)), dbl(sigma_scale))
    return sigmak, mu_self.value


# =========================================================================== Eliashberg
def linearized_eliashberg(chi, Gk, uni, init_delta, Smat, Cmat, olist, plist, kmap, invk, Nx, Ny, Nz, temp,
                          gap_sym, eps=1.0e-5, itemax=300, arnoldi_m=10):
    """(delta[No,No,Nw,Nk], lambda)  [_eliashberg.py:145-181]; chi is overwritten with V_Delta"""
    Norb, Nchi = len(Gk), len(Smat)
    Nkall, Nk, Nw = len(kmap), Gk.shape[3], Gk.shape[2]
    delta = np.zeros((Norb, Norb, Nw, Nk), dtype=np.complex128)
    lam = c_double(0.0)
    _call("lin_eliash", delta, chi, np.ascontiguousarray(Gk), np.ascontiguousarray(uni),
          np.ascontiguousarray(init_delta, dtype=np.float64), Smat, Cmat, olist,
          np.asarray(plist, dtype=np.float64), kmap, invk, dbl(temp), dbl(eps), i64(Nkall), i64(Nk), i64(Nw),
          i64(Nchi), i64(Norb), i64(Nx), i64(Ny), i64(Nz), i64(itemax), i64(gap_sym), i64(arnoldi_m), byref(lam))
    return delta, lam.value


def nonlinear_eliashberg(delta_init, Gk, hamk, Smat, Cmat, olist, plist, kmap, invk, mu, temp, gap_sym, Nx, Ny, Nz,
                         sw_sigma=False, sw_Vconst=False, eps=1.0e-4, itemax=100, m_diis=5, gap_min=1.0e-2,
                         sw_amp_newton=True):
    """(delta[No,No,Nw,Nk], sigmak[No,No,Nw,Nk])  [_eliashberg.py:230-304].
    delta_init is updated in place when it is C-contiguous complex128 (else a
    copy is returned); Gk, when C-contiguous complex128, leaves holding the last
    normal G, as the reference's intent(inout) argument does."""
    if delta_init.dtype != np.complex128 or not delta_init.flags['C_CONTIGUOUS']:
        delta = np.ascontiguousarray(delta_init, dtype=np.complex128)
    else:
        delta = delta_init
    Norb, _, Nw, Nk = delta.shape
    Nkall, Nchi = len(kmap), len(Smat)
    sigmak = np.zeros((Norb, Norb, Nw, Nk), dtype=np.complex128)
    G = Gk if (Gk.dtype == np.complex128 and Gk.flags['C_CONTIGUOUS']) else np.ascontiguousarray(Gk, np.complex128)
    _call("eliashberg", delta, sigmak, G, np.ascontiguousarray(hamk), Smat, Cmat, olist,
          np.asarray(plist, dtype=np.float64), kmap, invk, dbl(mu), dbl(temp), dbl(eps), i64(Nkall), i64(Nk), i64(Nw),
          i64(Nchi), i64(Norb), i64(Nx), i64(Ny), i64(Nz), i64(itemax), i64(gap_sym), byref(c_bool(sw_sigma)),
          byref(c_bool(sw_Vconst)), i64(m_diis), dbl(gap_min), byref(c_bool(sw_amp_newton)))
    return delta, sigmak


def eliashberg_stats():
    """Of the last nonlinear_eliashberg call: iterations, status ("itemax",
    "converged", "collapsed"), re-anchors of the amplitude phase, last relative
    change max|D_new - D| / max|D|."""
    out = np.zeros(4)
    _L().gmod_eliashberg_stats(out, 4)
    return dict(iterations=int(out[0]), status=("itemax", "converged", "collapsed")[int(out[1])],
                anchors=int(out[2]), rel_change=float(out[3]))


def get_chi0sc(Gk, fk, kmap, invk, olist, plist, temp, Nx, Ny, Nz, gap_sym):
    """(chi0_S, chi0_C)[Nchi,Nchi,Nw,Nk] = chi^GG +- chi^FF of the
    superconducting loop (ffeliash.f90:459-489)."""
    Nkall, Nk = len(kmap), Gk.shape[3]
    Norb, Nchi, Nw = len(Gk), len(olist), Gk.shape[2]
    cs = np.zeros((Nchi, Nchi, Nw, Nk), dtype=np.complex128)
    cc = np.zeros_like(cs)
    chi_map, irr_chi = _get_chi_map(olist)
    _call("get_chi0sc_", cs, cc, np.ascontiguousarray(Gk), np.ascontiguousarray(fk), kmap, invk, irr_chi, chi_map,
          olist, np.asarray(plist, dtype=np.float64), dbl(temp), i64(Nx), i64(Ny), i64(Nz), i64(Nw), i64(Nk),
          i64(Nkall), i64(Norb), i64(Nchi), byref(c_bool(gap_sym >= 0)))
    return cs, cc


def mkV_flex_nosoc(chiS, chiC, Smat, Cmat, gap_sym):
    """(V_sigma, V_Delta) of the superconducting loop from chi0_S, chi0_C
    (ffeliash.f90:386-457); inputs are not modified."""
    Nchi, _, Nw, Nk = chiS.shape
    vs = np.array(chiS, dtype=np.complex128, order='C')
    vd = np.array(chiC, dtype=np.complex128, order='C')
    _call("mkv_flex_nosoc_", vd, vs, Smat, Cmat, i64(Nk), i64(Nw), i64(Nchi), byref(c_bool(gap_sym >= 0)))
    return vs, vd


def gen_Fk(Gk, delta, invk):
    """F = -G Delta G^dagger  [_eliashberg.py:343-355]"""
    Nk, Nw, Norb = Gk.shape[3], delta.shape[2], len(delta)
    Fk = np.zeros((Norb, Norb, Nw, Nk), dtype=np.complex128)
    _call("mkfk_trs_nsoc_", Fk, np.ascontiguousarray(Gk), np.ascontiguousarray(delta), i64(Nk), i64(Nw), i64(Norb))
    return Fk


def get_initial_delta(init_delta, uni, kmap, invk, Nw, gap_sym):
    """[_eliashberg.py:378-399]"""
    Norb, Nk = np.shape(init_delta)
    delta = np.zeros((Norb, Norb, Nw, Nk), dtype=np.complex128)
    _call("get_initial_delta_", delta, np.ascontiguousarray(init_delta, dtype=np.float64),
          np.ascontiguousarray(uni), kmap, invk, i64(len(kmap)), i64(Nk), i64(Nw), i64(Norb), i64(gap_sym))
    return delta


def conv_delta_orb_to_band(delta, uni, invk, plist, gap_sym):
    """[_eliashberg.py:306-323]"""
    Nkall, Nk, Nw, Norb = len(invk), len(uni), delta.shape[2], len(delta)
    deltab = np.zeros((Norb, Norb, Nkall), dtype=np.complex128)
    _call("conv_delta_orb_to_band", deltab, np.ascontiguousarray(delta), np.ascontiguousarray(uni),
          np.asarray(plist, dtype=np.float64), invk, i64(Norb), i64(Nkall), i64(Nk), i64(Nw), i64(gap_sym))
    return deltab


def remap_gap(delta0, plist, invk, gap_sym):
    """[_eliashberg.py:426-441]"""
    Nkall, Nk, Norb = len(invk), delta0.shape[-1], len(plist)
    Nw = delta0.shape[2]
    delta = np.zeros((Norb, Norb, Nw, Nkall), dtype=np.complex128)
    _call("remap_delta_", delta, np.ascontiguousarray(delta0), np.asarray(plist, dtype=np.float64), invk,
          i64(Nkall), i64(Nk), i64(Nw), i64(Norb), i64(gap_sym))
    return delta


# =========================================================================== additions
def set_precision(mode):
    """0: FP64; 1: FP32 convolutions with FP64 judge (see include/gmod.h)."""
    _L().gmod_set_precision(int(mode))
    check()


def get_precision():
    return _L().gmod_get_precision()


def set_sigma_seed(sigmak, mu, mu_old):
    """In-memory warm start consumed by the next mkself(sw_in=True)."""
    s = np.ascontiguousarray(sigmak, dtype=np.complex128)
    _L().gmod_set_sigma_seed(s.ravel(), float(mu), float(mu_old), s.size)
    check()


def flex_stats():
    out = np.zeros(12)
    _L().gmod_flex_stats(out, 12)
    keys = ("iterations", "esterr", "stoner", "mu", "seconds", "t_chi0", "t_vertex", "t_sigma", "t_mu_dyson",
            "increment_iterations", "solve_seconds", "reference_moves")
    return dict(zip(keys, out.tolist()))


PROFILE_CLASSES = ("fft_pass", "rpa_vertex", "dyson", "scatter", "product", "gather", "fft_pass_fp32")


def profile(on=True):
    _L().gmod_profile(1 if on else 0)


def profile_read():
    out = np.zeros(3 * len(PROFILE_CLASSES))
    _L().gmod_profile_read(out, len(out))
    return {c: dict(ms=out[3 * i], launches=int(out[3 * i + 1]), bytes=out[3 * i + 2])
            for i, c in enumerate(PROFILE_CLASSES)}


def set_verbose(level):
    _L().gmod_set_verbose(int(level))


def pool_bytes():
    """Device memory held by the library's caching pool."""
    return float(_L().gmod_pool_bytes())


def release_memory():
    """Return the unused blocks of the library's device pool to HIP."""
    _L().gmod_release_memory()
    check()


class DeviceArray:
    """A device buffer of the library (gmod_device_alloc) holding `n` elements
    of numpy dtype `dtype`, for the device-pointer entry points.  It keeps the
    process on the library's own HIP runtime."""

    def __init__(self, n, dtype):
        self.dtype = np.dtype(dtype)
        self.n = int(n)
        self.nbytes = self.n * self.dtype.itemsize
        self.ptr = _L().gmod_device_alloc(max(self.nbytes, 1))
        check()
        if not self.ptr:
            raise RuntimeError("gmod_device_alloc failed")

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.size != self.n:
            raise ValueError("DeviceArray.upload: size mismatch")
        if _L().gmod_memcpy(self.ptr, a.ctypes.data, self.nbytes, 1) != 0:
            check()

    def download(self):
        out = np.empty(self.n, dtype=self.dtype)
        if _L().gmod_memcpy(out.ctypes.data, self.ptr, self.nbytes, 2) != 0:
            check()
        return out

    def free(self):
        if self.ptr:
            _L().gmod_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def fft_device(buf, tmp, dims, sign=-1, mask=15, reps=1):
    """In-place transform of a device grid (gmod_fft_device): buf and tmp are
    DeviceArray of complex64 (FP32 arithmetic) or complex128 (FP64) with
    nx*ny*nz*m elements, x fastest; dims = (nx, ny, nz, m)."""
    if buf.dtype not in (np.complex64, np.complex128) or tmp.dtype != buf.dtype:
        raise TypeError("fft_device: complex64 or complex128 buffers of one dtype")
    nx, ny, nz, m = (int(v) for v in dims)
    if buf.n != nx * ny * nz * m or tmp.n != buf.n:
        raise ValueError("fft_device: buffer size does not match dims")
    prec = 1 if buf.dtype == np.complex64 else 0
    rc = _L().gmod_fft_device(buf.ptr, tmp.ptr, nx, ny, nz, m, int(mask), int(sign), prec, int(reps))
    check()
    if rc != 0:
        raise RuntimeError("gmod_fft_device failed")


def regrid_sigma(sigmak_c, kmap_c, invk_c, mesh_c, kmap_f, invk_f, mesh_f, temp):
    """Sigma[No,No,Nw_c,Nk_c] of a coarse solve onto the fine mesh (same T):
    Fourier interpolation in k, shared Matsubara frequencies kept, tail
    c0 + c1/(i w) beyond (gmod_regrid_sigma).  mesh = (Nx, Ny, Nz, Nw).
    Returns Sigma[No,No,Nw_f,Nk_f]."""
    s = np.ascontiguousarray(sigmak_c, dtype=np.complex128)
    No, _, Nwc, Nkc = s.shape
    Nxc, Nyc, Nzc, Nwc2 = mesh_c
    Nxf, Nyf, Nzf, Nwf = mesh_f
    if Nwc2 != Nwc:
        raise ValueError("mesh_c Nw does not match sigmak_c")
    Nkf = int(invk_f[:, 0].max())
    out = np.zeros((No, No, Nwf, Nkf), dtype=np.complex128)
    _L().gmod_regrid_sigma(out, s, np.ascontiguousarray(kmap_c, dtype=np.int64),
                           np.ascontiguousarray(invk_c, dtype=np.int64), Nkc, Nxc, Nyc, Nzc, Nwc,
                           np.ascontiguousarray(kmap_f, dtype=np.int64), np.ascontiguousarray(invk_f, dtype=np.int64),
                           Nkf, Nxf, Nyf, Nzf, Nwf, No, float(temp))
    check()
    return out


def wannier_r(data, kmap, invk, Nx, Ny, Nz, N_cut, full=False):
    """data[Norb, Norb, Nw, Nk] on the irreducible k points (full: on every
    grid point, placed through kmap) -> data_R[Norb, Norb, min(N_cut, Nw), Nx,
    Ny, Nz], the normalised inverse transform over the k grid of the first
    N_cut frequencies (numpy's ifftn), computed by gmod_wannier_r on the
    device (the transform of _wannier_io.py:48-99)."""
    d = np.ascontiguousarray(data, dtype=np.complex128)
    No, _, Nw, Nk = d.shape
    Ncut = min(int(N_cut), Nw)
    out = np.zeros((No, No, Ncut, Nz, Ny, Nx), dtype=np.complex128)
    _L().gmod_wannier_r(out, d, np.ascontiguousarray(kmap, dtype=np.int64), np.ascontiguousarray(invk, dtype=np.int64),
                        Nx, Ny, Nz, Nw, Nk, No, Ncut, 1 if full else 0)
    check()
    return out.transpose(0, 1, 2, 5, 4, 3)
