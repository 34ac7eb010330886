"""The box-pass FFT of libgmod.so, checked on the CPU through its host hook
(gmod_box_fft_host runs the kernel's own load / sub-pass / store code on the
host with bounds checks) against numpy's FFT.  Covers the plan shapes of the
benchmark workloads (one spatial box + one t box in FP32, x-y / z / t boxes
in FP64), mixed radices 2-32 with 3, 5 and the four-step splits of long t
axes, and a small LDS budget that forces the split plans.

Conventions (libs/src/fmod.f90:67-160): forward sign -1 unnormalised,
backward sign +1 divided by the number of transformed points."""
import ctypes

import numpy as np
import pytest

from pyrpa_amd import _loader


def _hook():
    g = _lib_plan()

    def f(d, *rest):
        return g(d.ctypes.data_as(ctypes.c_void_p), *rest)
    return f


SHAPES = [(4, 4, 1, 32), (8, 8, 1, 320), (3, 3, 1, 10), (4, 3, 1, 8), (32, 32, 16, 16), (64, 64, 1, 128),
          (4, 4, 4, 8), (2, 2, 4, 8), (128, 128, 1, 8), (8, 4, 1, 2048), (16, 16, 1, 4096), (32, 8, 1, 2048),
          (16, 4, 2, 1024)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("prec", [0, 1])
def test_box_fft_matches_numpy(shape, prec):
    f = _hook()
    nx, ny, nz, m = shape
    rng = np.random.default_rng(sum(shape) + prec)
    a = rng.standard_normal((m, nz, ny, nx)) + 1j * rng.standard_normal((m, nz, ny, nx))
    tol = 1e-13 if prec == 0 else 3e-6
    covered = 0
    for mask in (15, 7):
        axes = [3 - i for i in range(4) if (mask >> i) & 1]
        for sign in (-1, 1):
            ref = np.fft.fftn(a, axes=axes) if sign < 0 else np.fft.ifftn(a, axes=axes)
            for lds in (163840, 16384):
                d = a.ravel().copy()
                npass = ctypes.c_int(0)
                rc = f(d, nx, ny, nz, m, mask, sign, prec, lds, ctypes.byref(npass))
                assert rc in (0, 1), (shape, mask, sign, lds, rc)
                if rc == 1:
                    continue
                covered += 1
                err = np.abs(d.reshape(a.shape) - ref).max() / np.abs(ref).max()
                assert err < tol, (shape, prec, mask, sign, lds, err)
    assert covered >= 4


def _plan_passes(shape, prec, mask=15, sign=-1, lds=163840):
    f = _lib_plan()
    npass = ctypes.c_int(0)
    assert f(None, *shape, mask, sign, prec, lds, ctypes.byref(npass)) == 0
    return npass.value


def _lib_plan():
    f = _loader.lib().gmod_box_fft_host
    f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                  ctypes.c_int, ctypes.c_long, ctypes.POINTER(ctypes.c_int)]
    f.restype = ctypes.c_int
    return f


@pytest.mark.parametrize("prec,passes", [(1, 2), (0, 3)])
def test_box_plan_pass_count_of_the_4orb_3d_mesh(prec, passes):
    """32x32x16 x 2048 with the 160 KB LDS of gfx950: FP32 = one spatial box
    (128 KB) + one t box (8 columns x 2048 slots); FP64 = x-y box + z box +
    t box."""
    assert _plan_passes((32, 32, 16, 2048), prec) == passes
    assert _plan_passes((32, 32, 16, 8), prec) == passes


@pytest.mark.parametrize("prec", [0, 1])
def test_split_t_axis_plan_matches_numpy(prec, monkeypatch):
    """The optional four-step split of the t axis (two strided-slot passes)."""
    monkeypatch.setenv("GMOD_FFT_TSPLIT", "1")
    f = _hook()
    for shape in [(32, 8, 1, 2048), (16, 16, 1, 4096)]:
        nx, ny, nz, m = shape
        assert _plan_passes(shape, prec) == (3 if prec == 1 else 3)
        rng = np.random.default_rng(m + prec)
        a = rng.standard_normal((m, nz, ny, nx)) + 1j * rng.standard_normal((m, nz, ny, nx))
        for sign in (-1, 1):
            d = a.ravel().copy()
            assert f(d, nx, ny, nz, m, 15, sign, prec, 163840, ctypes.byref(ctypes.c_int(0))) == 0
            ref = np.fft.fftn(a) if sign < 0 else np.fft.ifftn(a)
            err = np.abs(d.reshape(a.shape) - ref).max() / np.abs(ref).max()
            assert err < (1e-13 if prec == 0 else 3e-6), (shape, prec, sign, err)


@pytest.mark.parametrize("prec", [0, 1])
def test_spatial_first_backward_matches_numpy(prec):
    """The chi0 inverse transforms run the forward pass list (spatial axes
    first) with the backward twiddles (sign 2 of the hook).  A transform is
    separable, so the order must not change the result: against numpy's
    ifftn, on the benchmark's box shapes and on the split plans of a small LDS
    budget.  A pass list run with the wrong twiddle table or without the
    final 1/N gives an O(1) error here."""
    f = _hook()
    tol = 1e-13 if prec == 0 else 3e-6
    covered = 0
    for shape in [(32, 32, 16, 16), (8, 8, 1, 320), (16, 4, 2, 1024), (5, 3, 1, 42)]:
        nx, ny, nz, m = shape
        rng = np.random.default_rng(m + prec)
        a = rng.standard_normal((m, nz, ny, nx)) + 1j * rng.standard_normal((m, nz, ny, nx))
        for lds in (163840, 16384):
            d = a.ravel().copy()
            rc = f(d, nx, ny, nz, m, 15, 2, prec, lds, ctypes.byref(ctypes.c_int(0)))
            assert rc in (0, 1), (shape, lds, rc)
            if rc == 1:
                continue
            covered += 1
            ref = np.fft.ifftn(a)
            err = np.abs(d.reshape(a.shape) - ref).max() / np.abs(ref).max()
            assert err < tol, (shape, prec, lds, err)
    assert covered >= 4
