"""Adjoint transforms and convolutions on the GPU.

Adjoints are pinned by the dot-product identity <A x, y> = <x, A^H y>
(the reference's own check, src/matlab/so3_adjoint_forward_test.m:41-65 and
so3_adjoint_inverse_test.m:45-64), with A the GPU direct transform that
tests/test_gpu_transforms.py pins against the oracle.  The S^2 -> SO(3)
convolution is compared with the oracle, the SO(3) harmonic convolution
with the brute-force sum of tests/test_convolution.c:134-177 (tolerance
1e-12 as there).
"""
import ctypes

import numpy as np
import pytest

import so3_amd.api as A
import so3_amd.params as P
from so3_amd import synth
from so3_amd._lib import lib

pytestmark = pytest.mark.gpu


def _s(p):
    return ctypes.byref(p.struct())


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def adjoint_inverse(f, p):
    flmn = np.zeros(A.flmn_size(p), complex)
    fn = lib().so3_adjoint_inverse_direct_real if p.reality else lib().so3_adjoint_inverse_direct
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(P.so3_parameters_t)]
    fn(_ptr(flmn), _ptr(np.ascontiguousarray(f)), _s(p))
    return flmn


def adjoint_forward(flmn, p):
    q = p.replace(sampling_scheme=0)
    f = np.zeros(A.f_size(q), float if p.reality else complex)
    fn = lib().so3_adjoint_forward_direct_real if p.reality else lib().so3_adjoint_forward_direct
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(P.so3_parameters_t)]
    fn(_ptr(f), _ptr(np.ascontiguousarray(flmn)), _s(p))
    return f


def _rand_f(p, rng):
    n = A.f_size(p.replace(sampling_scheme=0))
    if p.reality:
        return rng.uniform(-1, 1, n)
    return rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)


def _dot(a, b):
    return np.vdot(a, b)  # conj(a) . b


@pytest.mark.parametrize("L,N,L0,real,storage,order,mode", [
    (8, 8, 0, 0, 0, 0, 0), (9, 4, 1, 0, 1, 1, 0), (8, 5, 0, 1, 0, 0, 0), (10, 6, 0, 0, 0, 1, 2),
    (12, 12, 0, 0, 1, 0, 0), (7, 3, 0, 1, 1, 0, 3), (16, 16, 0, 0, 0, 0, 0),
    # full size: the register-resident beta and DFT kernels (adjoint branch)
    (256, 256, 0, 0, 0, 0, 0), (256, 256, 0, 1, 0, 0, 0), (200, 150, 3, 0, 0, 1, 0),
    # L = 129..192: 1024-point convolution on the one-sequence register kernel
    (150, 40, 0, 0, 0, 0, 0), (170, 12, 0, 1, 0, 0, 0)])
def test_adjoint_dot_products(L, N, L0, real, storage, order, mode):
    p = P.SO3Parameters(L=L, N=N, L0=L0, reality=real, storage=storage, n_order=order, n_mode=mode)
    rng = np.random.default_rng(L + 100 * N)
    x = synth.gen_flmn(p, 3)        # coefficients
    y = _rand_f(p, rng)             # samples
    # inverse A: flmn -> f, adjoint A^H: f -> flmn
    Ax = A.inverse(x, p, A.DIRECT)
    AHy = adjoint_inverse(y, p)
    l, m, n = synth.nonzero_lmn(p)
    idx = P.elmn2ind_vec(l, m, n, p)
    lhs = _dot(Ax, y) if not real else np.dot(Ax, y)
    rhs = _dot(x[idx], AHy[idx])
    if real:
        # a real signal's inner product counts every n > 0 coefficient twice
        w = np.where(n > 0, 2.0, 1.0)
        rhs = np.sum(w * np.conj(x[idx]) * AHy[idx]).real
        lhs = lhs.real
    # relative to |lhs| (small cases) or to the Cauchy-Schwarz bound of the
    # inner product (large cases, where |lhs| is a sum of ~1e8 random terms)
    bound = np.linalg.norm(Ax) * np.linalg.norm(y)
    assert abs(lhs - rhs) <= max(1e-12 * max(1.0, abs(lhs)), 1e-13 * bound)
    # forward B: f -> flmn, adjoint B^H: flmn -> f
    Bf = A.forward(y, p, A.DIRECT)
    BHx = adjoint_forward(x, p)
    lhs = _dot(x[idx], Bf[idx])
    rhs = _dot(BHx, y) if not real else np.sum(BHx * y)
    if real:
        w = np.where(n > 0, 2.0, 1.0)
        lhs = np.sum(w * np.conj(x[idx]) * Bf[idx]).real
        rhs = rhs.real
    bound = np.linalg.norm(BHx) * np.linalg.norm(y)
    assert abs(lhs - rhs) <= max(1e-12 * max(1.0, abs(lhs)), 1e-13 * bound)


@pytest.mark.parametrize("L,N,real,storage,order", [(8, 8, 0, 0, 0), (16, 4, 0, 1, 1), (9, 3, 1, 0, 0), (33, 8, 0, 0, 1)])
def test_s2toso3_vs_oracle(oracle, L, N, real, storage, order):
    hp = P.SO3Parameters(L=L, N=N, reality=real, storage=storage, n_order=order)
    flm = synth.gen_flm(L, 1)
    glm = synth.gen_flm(L, 2)
    h = np.zeros(A.flmn_size(hp), complex)
    lib().so3_conv_s2toso3_harmonic_convolution.argtypes = [ctypes.c_void_p, ctypes.POINTER(P.so3_parameters_t),
                                                           ctypes.c_void_p, ctypes.c_void_p]
    lib().so3_conv_s2toso3_harmonic_convolution(_ptr(h), _s(hp), _ptr(flm), _ptr(glm))
    ref = oracle.s2toso3(flm, glm, hp)
    assert np.abs(h - ref).max() <= 1e-14 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("N", [8, 7])
@pytest.mark.parametrize("real", [0, 1])
def test_harmonic_convolution_brute_force(N, real):
    p = P.SO3Parameters(L=8, N=N, reality=real, n_order=0)
    lib().so3_conv_get_parameters_of_convolved_lmn.restype = P.so3_parameters_t
    lib().so3_conv_get_parameters_of_convolved_lmn.argtypes = [ctypes.POINTER(P.so3_parameters_t)] * 2
    hs = lib().so3_conv_get_parameters_of_convolved_lmn(_s(p), _s(p))
    hp = P.SO3Parameters(**{k: getattr(hs, k) for k, _ in P.so3_parameters_t._fields_})
    assert (hp.L, hp.N, hp.L0) == (8, 8, 0)
    f = synth.gen_flmn(p, 1)
    g = synth.gen_flmn(p, 2)
    h = np.zeros(A.flmn_size(hp), complex)
    lib().so3_conv_harmonic_convolution.argtypes = [ctypes.c_void_p, ctypes.POINTER(P.so3_parameters_t)] * 3
    lib().so3_conv_harmonic_convolution(_ptr(h), _s(hp), _ptr(f), _s(p), _ptr(g), _s(p))
    ns, ne, ni = P.n_loop_values(p)
    for i in range(A.flmn_size(hp)):
        el, m, n = A.ind2elmn(i, hp)
        want = 0
        for k in range(ns, ne + 1, ni):
            s = ctypes.byref(p.struct())
            if lib().so3_sampling_is_elmn_non_zero(el, n, k, s):
                want += f[A.elmn2ind(el, m, k, p)] * np.conj(g[A.elmn2ind(el, n, k, p)])
        assert abs(h[i] - want) < 1e-12
