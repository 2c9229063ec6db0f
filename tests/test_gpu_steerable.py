"""Steerable signals (steerable = 1; complex, via SSHT): N gamma samples at
pi g / N (reference so3_core.c:115-130 and :219-222 inverse, :309-328
forward).

Expected values come from the oracle's non-steerable transform and the
definition of the steerable gamma step, evaluated here in numpy:
  inverse: f_st(g) = sum_n F_n exp(i n pi g / N), g < N, where
    F_n = (1 / (2N-1)) sum_g' f_ns(g') exp(-2 pi i n g' / (2N-1)) are the
    per-n samples of the non-steerable inverse f_ns;
  forward: the steerable per-n samples
    fn_n = (2 pi / N) sum_{g<N} f(g) exp(-i n pi g / N) for n of the parity
    of N - 1, zero for the other n, are the non-steerable per-n samples of
    f'(g') = (1 / 2 pi) sum_n fn_n exp(2 pi i n g' / (2N-1)), so the
    steerable forward of f equals the non-steerable forward of f'.
The CPU test pins these two formulas against each other: a signal whose n
all have the parity of N - 1 goes round trip through them exactly (N
samples of exp(i n pi g / N) over N such n form an invertible DFT).
Tolerances as in test_gpu_transforms.py: 1e-12 relative to max|expected|.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import so3_amd.api as A
import so3_amd.params as P
from so3_amd import synth

TOL = 1e-12


def _rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(1.0, np.abs(b).max())


def _grid(p):
    ss = p.sampling_scheme == P.SO3_SAMPLING_MW_SS
    return (2 * p.L if ss else 2 * p.L - 1) * (p.L + 1 if ss else p.L)


def steer_inverse_ref(oracle, flmn, p):
    """f_st from the oracle's non-steerable inverse (the definition above)."""
    N, M = p.N, _grid(p)
    f_ns = oracle.inverse(flmn, p.replace(steerable=0)).reshape(2 * N - 1, M)
    F = np.fft.fft(f_ns, axis=0) / (2 * N - 1)  # row k holds F_n, n = k mod (2N-1)
    n = np.arange(-(N - 1), N)
    g = np.arange(N)
    E = np.exp(1j * np.pi * np.outer(g, n) / N)
    return (E @ F[n % (2 * N - 1)]).ravel()


def steer_forward_ref(oracle, f, p, out):
    """flmn_st as the non-steerable forward of the equivalent samples f'."""
    N, M = p.N, _grid(p)
    n = np.arange(-(N - 1), N)
    g = np.arange(N)
    fn = (2 * np.pi / N) * (np.exp(-1j * np.pi * np.outer(n, g) / N) @ np.asarray(f).reshape(N, M))
    fn[(n + N - 1) % 2 == 1] = 0.0
    gp = np.arange(2 * N - 1)
    fp = np.exp(2j * np.pi * np.outer(gp, n) / (2 * N - 1)) @ fn / (2 * np.pi)
    return oracle.forward(fp.ravel(), p.replace(steerable=0), out=out)


def parity_flmn(p, seed):
    """Seeded coefficients with every n of the other parity than N - 1 zeroed."""
    flmn = synth.gen_flmn(p.replace(steerable=0), seed)
    for i in range(flmn.size):
        _, _, n = A.ind2elmn(i, p)
        if (n + p.N - 1) % 2:
            flmn[i] = 0.0
    return flmn


CASES = [
    # L, N, L0, sampling, storage, order, n_mode
    (8, 8, 0, 0, 0, 0, 0),
    (8, 3, 0, 1, 0, 0, 0),
    (16, 5, 2, 0, 1, 1, 0),
    (10, 6, 0, 0, 0, 0, 1),
    (10, 7, 0, 1, 0, 1, 3),
    (9, 1, 0, 0, 0, 0, 0),
    (33, 16, 0, 0, 0, 0, 0),
    (48, 24, 0, 1, 0, 0, 0),
]


def _params(L, N, L0, ss, storage, order, mode):
    return P.SO3Parameters(L=L, N=N, L0=L0, sampling_scheme=ss, storage=storage, n_order=order, n_mode=mode,
                           steerable=1)


@pytest.mark.parametrize("L,N,L0,ss,storage,order,mode", [c for c in CASES if c[0] <= 16])
def test_steerable_reference_round_trip(oracle, L, N, L0, ss, storage, order, mode):
    """The two numpy formulas are each other's inverse on parity-restricted
    signals (CPU only: oracle plus numpy)."""
    p = _params(L, N, L0, ss, storage, order, mode)
    flmn = parity_flmn(p, 3 + L)
    f = steer_inverse_ref(oracle, flmn, p)
    assert f.size == A.f_size(p)
    back = steer_forward_ref(oracle, f, p, np.zeros(A.flmn_size(p), dtype=np.complex128))
    assert _rel(back, flmn) < 1e-11
    # a planted mistake (the other parity kept) breaks the round trip
    bad = synth.gen_flmn(p.replace(steerable=0), 3 + L)
    if N > 1 and mode == 0:
        f_bad = steer_inverse_ref(oracle, bad, p)
        back_bad = steer_forward_ref(oracle, f_bad, p, np.zeros(A.flmn_size(p), dtype=np.complex128))
        assert _rel(back_bad, bad) > 1e-3


@pytest.mark.gpu
@pytest.mark.parametrize("L,N,L0,ss,storage,order,mode", CASES)
def test_steerable_vs_oracle(oracle, L, N, L0, ss, storage, order, mode):
    p = _params(L, N, L0, ss, storage, order, mode)
    flmn = synth.gen_flmn(p.replace(steerable=0), 11 + L)
    f_gpu = A.inverse(flmn, p)
    assert f_gpu.size == A.f_size(p) == N * _grid(p)
    f_ref = steer_inverse_ref(oracle, flmn, p)
    assert _rel(f_gpu, f_ref) < TOL
    # forward of arbitrary samples
    rng = np.random.default_rng(L * 17 + N)
    f_any = rng.uniform(-1, 1, f_ref.shape) + 1j * rng.uniform(-1, 1, f_ref.shape)
    init = np.zeros(A.flmn_size(p), dtype=np.complex128)
    b_gpu = A.forward(f_any, p, flmn=init.copy())
    b_ref = steer_forward_ref(oracle, f_any, p, init.copy())
    assert _rel(b_gpu, b_ref) < TOL
    # round trip of a parity-restricted signal
    pf = parity_flmn(p, 5 + L)
    back = A.forward(A.inverse(pf, p), p, flmn=init.copy())
    assert _rel(back, pf) < 1e-11


@pytest.mark.gpu
def test_steerable_large_round_trip():
    """L = N = 256 MW: the steerable gamma at length 512 beside the
    register-path stages."""
    p = P.SO3Parameters(L=256, N=256, steerable=1, n_order=P.SO3_N_ORDER_ZERO_FIRST)
    flmn = synth.gen_flmn(p.replace(steerable=0), 21)
    n_of = np.empty(flmn.size, dtype=np.int64)
    L2 = p.L * p.L
    # padded zero-first storage: block k = 2n (n >= 0) or -2n-1 (n < 0)
    k = np.arange(flmn.size) // L2
    n_of[:] = np.where(k % 2 == 0, k // 2, -(k + 1) // 2)
    flmn[(n_of + p.N - 1) % 2 == 1] = 0.0
    f = A.inverse(flmn, p)
    assert f.size == p.N * _grid(p)
    back = A.forward(f, p)
    assert _rel(back, flmn) < 1e-11


def test_steerable_real_rejected():
    """Real steerable transforms are rejected before any device work (the
    reference's real via-SSHT code has its steerable branches commented
    out, so3_core.c:526-548, 743-813); the status entry point reports it."""
    code = ("import ctypes, so3_amd.api as A, so3_amd.params as P\n"
            "p = P.SO3Parameters(L=8, N=4, reality=1, steerable=1)\n"
            "rc = A.lib().so3x_inverse(None, None, ctypes.byref(p.struct()), 0, None)\n"
            "print(rc, A.lib().so3x_last_error().decode())\n")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-c", code], cwd=root, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0
    assert r.stdout.startswith("-1 steerable real signals are not supported")
