"""GPU parity of the batched FFT entry points (g3d_fft_plan/exec, the FftMany drop-in of src/fftmany.hpp:26-40)
against the reference's own FftMany (CPU, src/fft.hpp radix 2) and the direct DFT on small sizes."""
import ctypes as C

import numpy as np
import pytest

import g3d_amd
import refpy

pytestmark = pytest.mark.gpu


def run_gpu(nx, ny, two_d, data, inverse):
    L = g3d_amd.lib()
    a = np.ascontiguousarray(data, dtype=np.complex128)
    d = C.c_void_p()
    g3d_amd.check(L.g3d_malloc(C.byref(d), a.nbytes))
    try:
        g3d_amd.check(L.g3d_memcpy_h2d(d, a.ctypes.data_as(C.c_void_p), a.nbytes))
        plan = C.c_void_p()
        g3d_amd.check(L.g3d_fft_plan(nx, ny, int(two_d), C.byref(plan)))
        g3d_amd.check(L.g3d_fft_exec(plan, d, a.shape[0], int(inverse)))
        g3d_amd.check(L.g3d_fft_destroy(plan))
        out = np.empty_like(a)
        g3d_amd.check(L.g3d_memcpy_d2h(out.ctypes.data_as(C.c_void_p), d, a.nbytes))
    finally:
        g3d_amd.check(L.g3d_free(d))
    return out


@pytest.mark.parametrize("nx,ny,count", [(4, 4, 3), (16, 8, 5), (64, 64, 7), (256, 256, 3), (128, 512, 2)])
@pytest.mark.parametrize("inverse", [0, 1])
def test_fft_2d(nx, ny, count, inverse):
    rng = np.random.default_rng(nx + ny + count)
    x = rng.standard_normal((count, ny, nx)) + 1j * rng.standard_normal((count, ny, nx))
    g = run_gpu(nx, ny, True, x, inverse)
    ref = np.fft.ifft2(x) if inverse else np.fft.fft2(x)        # forward sign -1, inverse scaled (fftmany.hpp:3)
    assert np.abs(g - ref).max() <= 1e-13 * np.abs(ref).max()
    if refpy.available():
        r = refpy.fft_many(nx, ny, True, x.copy(), inverse)
        assert np.abs(g - r).max() <= 1e-14 * np.abs(r).max()


@pytest.mark.parametrize("nx,ny,count", [(8, 6, 2), (512, 384, 2)])
def test_fft_rows(nx, ny, count):
    rng = np.random.default_rng(1)
    x = rng.standard_normal((count, ny, nx)) + 1j * rng.standard_normal((count, ny, nx))
    for inverse in (0, 1):
        g = run_gpu(nx, ny, False, x, inverse)
        ref = np.fft.ifft(x, axis=2) if inverse else np.fft.fft(x, axis=2)
        assert np.abs(g - ref).max() <= 1e-13 * np.abs(ref).max()


def test_fft_round_trip():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((4, 128, 128)) + 1j * rng.standard_normal((4, 128, 128))
    y = run_gpu(128, 128, True, run_gpu(128, 128, True, x, 0), 1)
    assert np.abs(y - x).max() <= 1e-14 * np.abs(x).max()
