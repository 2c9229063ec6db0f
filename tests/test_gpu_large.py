"""Full-size (L = N = 256 and 129 <= L, N <= 256) transforms on the GPU
against brute-force synthesis.

At this size the DFT stages run their register-resident 1024-point paths
(Bluestein inner length of 2L - 1 = 511), which the small parity cases do not
reach.  The oracle cannot run a full L = 256 transform in test time, so the
inverse is checked on a sparse coefficient set at random grid points against
the defining sum f = sum (2l+1)/(8 pi^2) f_lmn e^{i m alpha} d^l_mn(beta)
e^{i n gamma}, with d^l_mn(beta) = i^{n-m} sum_m' D_m'm D_m'n e^{i m' beta}
built from the oracle's Delta tables (that formula is itself checked against
the explicit Wigner sum at small l below); the forward transform is checked
by the round trip on the same input.
"""
import math

import numpy as np
import pytest

import so3_amd.api as so3
from so3_amd import params as P
from tests.test_oracle import _full, d_exact

pytestmark = pytest.mark.gpu


def _d_from_delta(D, l, m, n, beta):
    mp = np.arange(-l, l + 1)
    s = np.sum(D[:, m + l] * D[:, n + l] * np.exp(1j * mp * beta))
    return (1j ** ((n - m) % 4)) * s


def test_d_from_delta_formula(oracle):
    T = oracle.dl_quarter(8, 0)
    for l, m, n, beta in [(5, 2, -3, 0.7), (7, -7, 4, 2.1), (6, 0, 0, 1.3), (3, 1, 1, 0.2)]:
        D = _full(T[l], l)
        assert abs(_d_from_delta(D, l, m, n, beta) - d_exact(l, m, n, beta)) < 1e-13


COEFS = [(255, 255, -255), (255, -3, 254), (200, -17, 33), (128, 0, 0), (97, 96, -50),
         (1, 1, -1), (254, -200, 0), (60, 5, 60)]
REAL_COEFS = [(255, 255, 255), (255, -3, 254), (200, -17, 33), (128, 0, 0), (97, 96, 50),
              (1, 1, 1), (254, -200, 7), (60, 5, 60)]


def _fit(coefs, L, N, real):
    """The coefficient list mapped into 0 <= l < L, |m| <= l, |n| <= min(l, N-1)."""
    out = []
    for (l, m, n) in coefs:
        l = min(l, L - 1)
        m = max(-l, min(l, m))
        nn = min(abs(n), l, N - 1)
        n = nn if (real or n >= 0) else -nn
        if (l, m, n) not in out:
            out.append((l, m, n))
    return out


# L = N = 256 and sizes with every Bluestein length between 257 and 512 on the
# register FFT: alpha and beta of 2L - 1 (MW) or 2L (MW_SS), gamma of 2N - 1;
# L = 300 runs alpha and the inverse beta on the 2048-point register FFT and
# gamma between X2 and f (complex: the gamma tile kernel; real: two columns
# per DFT)
@pytest.mark.parametrize("L,N,real,ss", [(256, 256, 0, 0), (256, 256, 1, 0), (256, 256, 0, 1), (256, 256, 1, 1),
                                         (150, 140, 0, 0), (170, 131, 0, 1), (129, 129, 0, 0),
                                         (211, 180, 1, 0), (300, 140, 0, 0), (300, 150, 1, 0)])
def test_sparse_inverse_and_round_trip(oracle, L, N, real, ss):
    p = P.SO3Parameters(L=L, N=N, reality=real, sampling_scheme=ss, n_order=P.SO3_N_ORDER_ZERO_FIRST)
    rng = np.random.default_rng(7 + real + L)
    flmn = np.zeros(P.flmn_size(p), dtype=np.complex128)
    coefs = _fit(REAL_COEFS if real else COEFS, L, N, real)
    vals = {}
    for (l, m, n) in coefs:
        v = rng.uniform(-1, 1) + 1j * rng.uniform(-1, 1)
        if real and n == 0 and m == 0:
            v = v.real + 0j
        vals[(l, m, n)] = v
        flmn[P.elmn2ind_vec(np.array([l]), np.array([m]), np.array([n]), p)[0]] = v
    f = so3.inverse(flmn, p)
    na, nb, ng = (2 * L, L + 1, 2 * N - 1) if ss else (2 * L - 1, L, 2 * N - 1)
    assert f.size == na * nb * ng
    T = oracle.dl_quarter(L, 0)
    Ds = {l: _full(T[l], l) for l in {c[0] for c in coefs}}
    pts = rng.integers(0, [na, nb, ng], size=(300, 3))
    fmax = 0.0
    err = 0.0
    for a, b, g in pts:
        be = np.pi * b / L if ss else (2 * b + 1) * np.pi / (2 * L - 1)
        al, ga = 2 * np.pi * a / na, 2 * np.pi * g / ng
        ref = 0j
        for (l, m, n), c in vals.items():
            t = c * (2 * l + 1) / (8 * np.pi ** 2) * np.exp(1j * m * al) * \
                _d_from_delta(Ds[l], l, m, n, be) * np.exp(1j * n * ga)
            ref += t + np.conj(t) if (real and n > 0) else t
        if real:
            ref = ref.real
        got = f[(g * nb + b) * na + a]
        err = max(err, abs(got - ref))
        fmax = max(fmax, abs(ref))
    assert err <= 1e-12 * max(fmax, 1.0), (err, fmax)
    back = so3.forward(f, p)
    assert np.abs(back - flmn).max() < 1e-11


# Forward transforms of arbitrary (not band-limited) samples at 100 <= L <=
# 256, where the beta stages pair sequences of opposite symmetry and the
# pole samples of the antisymmetric member take a separate path: against the
# oracle's per-sequence algorithm (N small keeps the oracle fast).
@pytest.mark.parametrize("L,N,real,ss", [(256, 2, 0, 0), (256, 2, 1, 0), (256, 2, 0, 1), (256, 2, 1, 1),
                                         (200, 3, 0, 0), (193, 2, 0, 1),
                                         # 1024-point convolution on the register kernel
                                         (150, 2, 0, 0), (129, 3, 1, 1),
                                         # workgroup engine, 512-point direct convolution
                                         (120, 3, 0, 0), (100, 2, 1, 1),
                                         # alpha and the inverse beta on the 2048-point
                                         # register FFT (Pb = 513 ... 1024)
                                         (257, 3, 0, 0), (300, 2, 1, 0), (384, 2, 0, 1),
                                         (511, 1, 1, 1), (512, 2, 0, 0),
                                         # alpha and the inverse beta on the 4096-point
                                         # register FFT (Pb = 1025 ... 2048)
                                         (513, 1, 0, 0), (700, 1, 1, 1), (640, 2, 0, 1)])
def test_forward_arbitrary_samples_paired_beta(oracle, L, N, real, ss):
    p = P.SO3Parameters(L=L, N=N, reality=real, sampling_scheme=ss, n_order=P.SO3_N_ORDER_ZERO_FIRST)
    rng = np.random.default_rng(L + 10 * real + 100 * ss)
    n = so3.f_size(p)
    f = rng.uniform(-1, 1, n)
    if not real:
        f = f + 1j * rng.uniform(-1, 1, n)
    init = np.zeros(P.flmn_size(p), dtype=np.complex128)
    got = so3.forward(f, p, flmn=init.copy())
    ref = oracle.forward(f, p, "via_ssht", out=init.copy())
    assert np.abs(got - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())
    # the inverse of the same shape, through the paired inverse beta stage
    fi = so3.inverse(ref, p)
    fr = oracle.inverse(ref, p, "via_ssht")
    assert np.abs(fi - fr).max() <= 1e-12 * max(1.0, np.abs(fr).max())


# shapes of the batched workload's largest items (Bluestein inner lengths 2048,
# Pw = 1536, N = 1) and a long-alpha case, checked by the round trip
@pytest.mark.parametrize("L,N,ss,real", [(384, 64, 0, 0), (384, 64, 1, 1), (368, 17, 0, 1),
                                         (320, 1, 1, 0), (512, 4, 0, 0), (32, 32, 1, 0),
                                         (256, 256, 1, 0), (200, 200, 1, 1), (129, 150, 0, 0),
                                         (1024, 4, 0, 0)])
def test_round_trip_batch_shapes(L, N, ss, real):
    from so3_amd import synth
    p = P.SO3Parameters(L=L, N=N, sampling_scheme=ss, reality=real, n_order=P.SO3_N_ORDER_ZERO_FIRST)
    flmn = synth.gen_flmn(p, 5)
    back = so3.forward(so3.inverse(flmn, p), p)
    assert np.abs(back - flmn).max() < 1e-10 * max(1.0, np.abs(flmn).max())
