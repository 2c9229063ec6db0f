"""Adjoint transforms and convolutions on the GPU.

Adjoints are compared with the oracle's restatement of src/c/so3_adjoint.c
(itself pinned against brute-force matrices in tests/test_oracle.py) on
identical inputs, and at full size by the dot-product identity
<A x, y> = <x, A^H y> (the reference's own check,
src/matlab/so3_adjoint_forward_test.m:41-65 and so3_adjoint_inverse_test.m:45-64),
with A the GPU direct transform that tests/test_gpu_transforms.py pins against
the oracle.  The S^2 -> SO(3) convolution is compared with the oracle, the
SO(3) harmonic convolution with the brute-force sum of
tests/test_convolution.c:134-177 (tolerance 1e-12 as there), and the
real-space convolution with the oracle's direct forward -> harmonic product
-> direct inverse chain (src/c/so3_conv.c:154-184).
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


@pytest.mark.parametrize("L,N,L0,real,storage,order,mode", [
    (8, 8, 0, 0, 0, 0, 0), (9, 4, 1, 0, 1, 1, 0), (8, 5, 0, 1, 0, 0, 0), (10, 6, 0, 0, 0, 1, 2),
    (12, 12, 0, 1, 1, 0, 1), (7, 3, 0, 0, 1, 0, 3), (6, 4, 0, 0, 0, 0, 4), (16, 16, 2, 0, 0, 0, 0),
    (33, 9, 0, 1, 0, 1, 0)])
def test_adjoints_vs_oracle(oracle, L, N, L0, real, storage, order, mode):
    """GPU so3_adjoint_* against the oracle's so3_adjoint.c restatement on the
    same inputs (every stored coefficient, every sample), 1e-12 relative."""
    p = P.SO3Parameters(L=L, N=N, L0=L0, reality=real, storage=storage, n_order=order, n_mode=mode)
    rng = np.random.default_rng(L + 10 * N + mode)
    y = _rand_f(p, rng)
    got = adjoint_inverse(y, p)
    ref = oracle.adjoint_inverse(y, p)
    assert np.abs(got - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())
    x = synth.gen_flmn(p, 4)
    got = adjoint_forward(x, p)
    ref = oracle.adjoint_forward(x, p)
    assert np.abs(got - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())


def _harmonic_conv_ref(f, fp, g, gp, hp):
    """so3_conv_harmonic_convolution (so3_conv.c:44-90) as a brute-force sum"""
    h = np.zeros(A.flmn_size(hp), complex)
    ns, ne, ni = P.n_loop_values(fp)
    s = ctypes.byref(gp.struct())
    for i in range(A.flmn_size(hp)):
        el, m, n = A.ind2elmn(i, hp)
        if abs(m) > el or abs(n) > el or el < hp.L0:
            continue
        for k in range(ns, ne + 1, ni):
            if lib().so3_sampling_is_elmn_non_zero(el, n, k, s):
                h[i] += f[A.elmn2ind(el, m, k, fp)] * np.conj(g[A.elmn2ind(el, n, k, gp)])
    return h


@pytest.mark.parametrize("real_flag", [0, 1])
def test_convolution_vs_oracle_chain(oracle, real_flag):
    """so3_conv_convolution(h, f, g): direct forward of f and g, harmonic
    product, direct inverse of h (so3_conv.c:154-184), against the same chain
    on the oracle.  f, g and h are complex signals whatever `reality` says
    (the reference calls the complex direct transforms)."""
    fp = P.SO3Parameters(L=8, N=4, L0=1, reality=real_flag, n_order=0)
    gp = P.SO3Parameters(L=9, N=5, reality=real_flag, n_order=0)
    lib().so3_conv_get_parameters_of_convolved_lmn.restype = P.so3_parameters_t
    lib().so3_conv_get_parameters_of_convolved_lmn.argtypes = [ctypes.POINTER(P.so3_parameters_t)] * 2
    hs = lib().so3_conv_get_parameters_of_convolved_lmn(_s(fp), _s(gp))
    hp = P.SO3Parameters(**{k: getattr(hs, k) for k, _ in P.so3_parameters_t._fields_})
    assert (hp.L, hp.L0, hp.N) == (8, 1, 8)
    fc, gc, hc = fp.replace(reality=0), gp.replace(reality=0), hp.replace(reality=0)
    rng = np.random.default_rng(12)
    f = rng.uniform(-1, 1, A.f_size(fc)) + 1j * rng.uniform(-1, 1, A.f_size(fc))
    g = rng.uniform(-1, 1, A.f_size(gc)) + 1j * rng.uniform(-1, 1, A.f_size(gc))
    h = np.zeros(A.f_size(hc), complex)
    lib().so3_conv_convolution.argtypes = [ctypes.c_void_p, ctypes.POINTER(P.so3_parameters_t)] * 3
    lib().so3_conv_convolution(_ptr(h), _s(hp), _ptr(f), _s(fp), _ptr(g), _s(gp))
    fl = oracle.forward(f, fc, "direct")
    gl = oracle.forward(g, gc, "direct")
    hl = _harmonic_conv_ref(fl, fc, gl, gc, hc)
    ref = oracle.inverse(hl, hc, "direct")
    assert np.abs(h - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())


def test_complex_inverse_ignores_reality(oracle):
    """so3_core_inverse_via_ssht is the complex transform whatever `reality`
    says (the reference's complex path never reads it)."""
    p = P.SO3Parameters(L=8, N=5, n_order=0)
    flmn = synth.gen_flmn(p, 6)
    f = np.zeros(A.f_size(p), complex)
    pr = p.replace(reality=1)
    lib().so3_core_inverse_via_ssht(_ptr(f), _ptr(flmn), _s(pr))
    ref = oracle.inverse(flmn, p)
    assert np.abs(f - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())
