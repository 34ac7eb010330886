"""gmod_eig_or_tr_chi_noalloc: the allocation-free variant of get_eig_or_tr_chi
(libs/src/fself.f90:818-850) on device buffers, against the existing entry
point, on two calls in a row (the second smaller), for the leading
eigenvalue and the trace; a batch beyond its fixed workspace is refused with
an error and leaves the library working.  The eigenvalue kernel and the
k-map are the same code as the existing entry point's, so the results must
agree bit for bit."""
import numpy as np
import pytest

import oracle.pyrpa_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def F():
    import pyrpa_amd.flibs as F
    return F


def _case(nx, ny, nchi, seed):
    klist, kmap, invk = O.gen_irr_k_TRS(nx, ny, 1)
    rng = np.random.default_rng(seed)
    nk = len(klist)
    chi = 0.3 * (rng.standard_normal((nchi, nchi, nk)) + 1j * rng.standard_normal((nchi, nchi, nk)))
    return chi, invk


def _noalloc(F, chi, invk, sw):
    nchi, _, nk = chi.shape
    nkall = len(invk)
    d_chi = F.DeviceArray.from_host(chi.ravel())
    d_invk = F.DeviceArray.from_host(np.ascontiguousarray(invk).ravel())
    d_out = F.DeviceArray.from_host(np.full(nkall, np.nan + 1j * np.nan))
    rc = F._L().gmod_eig_or_tr_chi_noalloc(d_out.ptr, d_chi.ptr, d_invk.ptr, nkall, nk, nchi, 1 if sw else 0)
    return rc, d_out.download()


@pytest.mark.parametrize("sw", [True, False])
def test_noalloc_matches_the_allocating_entry_point_on_two_calls(F, sw):
    for mesh, nchi, seed in (((24, 24), 16, 1), ((8, 8), 4, 2)):
        chi, invk = _case(*mesh, nchi, seed)
        rc, got = _noalloc(F, chi, invk, sw)
        assert rc == 0
        want = F.get_eig_or_tr_chi(chi, invk, sw)
        assert np.array_equal(got, want), (mesh, nchi, np.abs(got - want).max())


def test_noalloc_refuses_a_batch_beyond_its_workspace(F):
    from pyrpa_amd import _loader
    chi, invk = _case(8, 8, 4, 3)
    nk = chi.shape[2]
    d_chi = F.DeviceArray.from_host(chi.ravel())
    d_invk = F.DeviceArray.from_host(np.ascontiguousarray(invk).ravel())
    d_out = F.DeviceArray(len(invk), np.complex128)
    # Nk so large that the interleaved Hessenberg workspace exceeds the fixed
    # array: refused before any kernel runs (the buffers are never read)
    rc = F._L().gmod_eig_or_tr_chi_noalloc(d_out.ptr, d_chi.ptr, d_invk.ptr, len(invk), 1 << 22, 25, 1)
    assert rc == 1
    with pytest.raises(RuntimeError, match="maximum"):
        _loader.check()
    rc, got = _noalloc(F, chi, invk, True)
    assert rc == 0 and np.array_equal(got, F.get_eig_or_tr_chi(chi, invk, True))
    assert nk < (1 << 22)
