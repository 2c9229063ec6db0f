"""The library's batched DFT kernel (so3x_dft) against numpy.fft."""
import ctypes

import numpy as np
import pytest

import so3_amd.api as A
from so3_amd._lib import lib

pytestmark = pytest.mark.gpu

LENGTHS = [1, 2, 3, 4, 5, 7, 8, 9, 11, 13, 15, 16, 17, 31, 61, 64, 73, 127, 255, 511, 512, 600,
           767, 1021, 1023, 1024, 1537, 2047, 4096]


def _hip():
    h = ctypes.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipDeviceSynchronize.argtypes = []
    return h


@pytest.mark.parametrize("n", LENGTHS)
@pytest.mark.parametrize("direction", [-1, 1])
def test_dft(n, direction):
    lib()
    h = _hip()
    rng = np.random.default_rng(n)
    howmany = 5
    x = rng.standard_normal((howmany, n)) + 1j * rng.standard_normal((howmany, n))
    x = np.ascontiguousarray(x)
    d = ctypes.c_void_p()
    assert h.hipMalloc(ctypes.byref(d), x.nbytes) == 0
    try:
        assert h.hipMemcpy(d, x.ctypes.data, x.nbytes, 1) == 0
        A.dft_device(d.value, n, howmany, direction)
        assert h.hipDeviceSynchronize() == 0
        y = np.empty_like(x)
        assert h.hipMemcpy(y.ctypes.data, d, x.nbytes, 2) == 0
    finally:
        h.hipFree(d)
    ref = np.fft.fft(x, axis=1) if direction < 0 else np.fft.ifft(x, axis=1) * n
    err = np.abs(y - ref).max() / np.abs(ref).max()
    assert err < 5e-15 * max(1.0, np.log2(n)), err


# Lengths with a Bluestein inner length of 1024 (257 <= n <= 512) run on the
# two-sequence register FFT, two rows per wavefront: odd and even n (the
# chirp's half table has sign (-1)^n), and row counts that leave the last
# wavefront with one row (1, 17) or a whole group (2).
@pytest.mark.parametrize("n", [257, 258, 300, 383, 384, 401, 447, 500, 509, 510])
@pytest.mark.parametrize("howmany", [1, 2, 17])
@pytest.mark.parametrize("direction", [-1, 1])
def test_dft_register_lengths(n, howmany, direction):
    lib()
    h = _hip()
    rng = np.random.default_rng(1000 * n + howmany)
    x = rng.standard_normal((howmany, n)) + 1j * rng.standard_normal((howmany, n))
    x = np.ascontiguousarray(x)
    d = ctypes.c_void_p()
    assert h.hipMalloc(ctypes.byref(d), x.nbytes) == 0
    try:
        assert h.hipMemcpy(d, x.ctypes.data, x.nbytes, 1) == 0
        A.dft_device(d.value, n, howmany, direction)
        assert h.hipDeviceSynchronize() == 0
        y = np.empty_like(x)
        assert h.hipMemcpy(y.ctypes.data, d, x.nbytes, 2) == 0
    finally:
        h.hipFree(d)
    ref = np.fft.fft(x, axis=1) if direction < 0 else np.fft.ifft(x, axis=1) * n
    err = np.abs(y - ref).max() / np.abs(ref).max()
    assert err < 5e-15 * np.log2(n), err


# Lengths with a Bluestein inner length of 2048 (513 <= n <= 1024) run on the
# 2048-point register FFT, one row per wavefront (eight per workgroup): odd
# and even n around the chirp table's halves (j = 512 reads c_{n-512}), and
# row counts below, at and past one workgroup.
@pytest.mark.parametrize("n", [513, 514, 600, 767, 768, 769, 1000, 1023, 1024])
@pytest.mark.parametrize("howmany", [1, 8, 9])
@pytest.mark.parametrize("direction", [-1, 1])
def test_dft_register_2048(n, howmany, direction):
    lib()
    h = _hip()
    rng = np.random.default_rng(7000 * n + howmany)
    x = rng.standard_normal((howmany, n)) + 1j * rng.standard_normal((howmany, n))
    x = np.ascontiguousarray(x)
    d = ctypes.c_void_p()
    assert h.hipMalloc(ctypes.byref(d), x.nbytes) == 0
    try:
        assert h.hipMemcpy(d, x.ctypes.data, x.nbytes, 1) == 0
        A.dft_device(d.value, n, howmany, direction)
        assert h.hipDeviceSynchronize() == 0
        y = np.empty_like(x)
        assert h.hipMemcpy(y.ctypes.data, d, x.nbytes, 2) == 0
    finally:
        h.hipFree(d)
    ref = np.fft.fft(x, axis=1) if direction < 0 else np.fft.ifft(x, axis=1) * n
    err = np.abs(y - ref).max() / np.abs(ref).max()
    assert err < 5e-15 * np.log2(n), err


# Lengths with a Bluestein inner length of 4096 (1025 <= n <= 2048) run on the
# 4096-point register FFT, one row per pair of wavefronts (four per
# workgroup): odd and even n, and row counts that leave the last workgroup
# with idle pairs (1, 5) or none (4).
@pytest.mark.parametrize("n", [1025, 1026, 1500, 1537, 2047, 2048])
@pytest.mark.parametrize("howmany", [1, 4, 5])
@pytest.mark.parametrize("direction", [-1, 1])
def test_dft_register_4096(n, howmany, direction):
    lib()
    h = _hip()
    rng = np.random.default_rng(9000 * n + howmany)
    x = rng.standard_normal((howmany, n)) + 1j * rng.standard_normal((howmany, n))
    x = np.ascontiguousarray(x)
    d = ctypes.c_void_p()
    assert h.hipMalloc(ctypes.byref(d), x.nbytes) == 0
    try:
        assert h.hipMemcpy(d, x.ctypes.data, x.nbytes, 1) == 0
        A.dft_device(d.value, n, howmany, direction)
        assert h.hipDeviceSynchronize() == 0
        y = np.empty_like(x)
        assert h.hipMemcpy(y.ctypes.data, d, x.nbytes, 2) == 0
    finally:
        h.hipFree(d)
    ref = np.fft.fft(x, axis=1) if direction < 0 else np.fft.ifft(x, axis=1) * n
    err = np.abs(y - ref).max() / np.abs(ref).max()
    assert err < 5e-15 * np.log2(n), err
