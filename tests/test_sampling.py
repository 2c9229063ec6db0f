"""Sampling and index helpers of the C ABI (bit-exact) against the
reference's known-answer tables, its closed forms and round-trip
invariants; the oracle's and the Python generator's index maps against the
library's."""
import ctypes
import json
import math
import os

import numpy as np
import pytest

import so3_amd.api as A
import so3_amd.params as P
from so3_amd._lib import lib

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "index_kat.json")))
KAT_KEYS = [k for k in GOLD if not k.startswith("_") and k != "convolved_parameters"]


@pytest.mark.parametrize("key", KAT_KEYS)
def test_index_known_answers(key):
    t = GOLD[key]
    p = P.SO3Parameters(L=t["L"], N=t["N"], storage=t["storage"], n_order=t["n_order"],
                        reality=t["reality"])
    for el, m, n, ind in t["cases"]:
        assert A.elmn2ind(el, m, n, p) == ind
        assert A.ind2elmn(ind, p) == (el, m, n)


def _all_lmn(p):
    out = []
    ns = range(0, p.N) if p.reality else range(-(p.N - 1), p.N)
    for n in ns:
        lo = abs(n) if p.storage == P.SO3_STORAGE_COMPACT else 0
        for el in range(lo, p.L):
            for m in range(-el, el + 1):
                out.append((el, m, n))
    return out


@pytest.mark.parametrize("storage", [0, 1])
@pytest.mark.parametrize("order", [0, 1])
@pytest.mark.parametrize("real", [0, 1])
@pytest.mark.parametrize("L,N", [(5, 3), (6, 6), (4, 1)])
def test_index_bijection(storage, order, real, L, N):
    """elmn2ind is a bijection onto [0, flmn_size) and ind2elmn its inverse;
    the vectorised Python map used by the generator agrees."""
    p = P.SO3Parameters(L=L, N=N, storage=storage, n_order=order, reality=real)
    size = A.flmn_size(p)
    lmn = _all_lmn(p)
    idx = [A.elmn2ind(el, m, n, p) for el, m, n in lmn]
    assert sorted(idx) == list(range(size))
    for (el, m, n), i in zip(lmn, idx):
        assert A.ind2elmn(i, p) == (el, m, n)
    l, m, n = (np.array(v) for v in zip(*lmn))
    assert np.array_equal(P.elmn2ind_vec(l, m, n, p), np.array(idx))


def test_oracle_index_map_matches(oracle):
    lib_o = oracle.lib()
    lib_o.oracle_elmn2ind.argtypes = [ctypes.c_int] * 3 + [ctypes.POINTER(oracle.oracle_params)]
    lib_o.oracle_elmn2ind_real.argtypes = [ctypes.c_int] * 3 + [ctypes.POINTER(oracle.oracle_params)]
    for storage in (0, 1):
        for order in (0, 1):
            for real in (0, 1):
                p = P.SO3Parameters(L=7, N=4, storage=storage, n_order=order, reality=real)
                q = oracle._params(p)
                fn = lib_o.oracle_elmn2ind_real if real else lib_o.oracle_elmn2ind
                for el, m, n in _all_lmn(p):
                    if storage == 1 or abs(n) <= el:
                        assert fn(el, m, n, ctypes.byref(q)) == A.elmn2ind(el, m, n, p)


@pytest.mark.parametrize("L,N,steer", [(4, 3, 0), (7, 7, 0), (5, 2, 1)])
@pytest.mark.parametrize("ss", [0, 1])
def test_sizes_closed_forms(L, N, steer, ss):
    """so3_sampling.c:80-206,316-339"""
    p = P.SO3Parameters(L=L, N=N, sampling_scheme=ss, steerable=steer)
    na, nb = (2 * L - 1, L) if ss == 0 else (2 * L, L + 1)
    ng = N if steer else 2 * N - 1
    assert (A.n_alpha(p), A.n_beta(p), A.n_gamma(p)) == (na, nb, ng)
    assert A.f_size(p) == na * nb * ng
    n_expected = ((2 * L - 1) * (L - 1) + 1) * ng if ss == 0 else ((2 * L) * (L - 1) + 2) * ng
    assert lib().so3_sampling_n(ctypes.byref(p.struct())) == n_expected
    for storage, real, want in [(0, 0, (2 * N - 1) * L * L), (0, 1, N * L * L),
                                (1, 0, (2 * N - 1) * (3 * L * L - N * (N - 1)) // 3),
                                (1, 1, N * (6 * L * L - (N - 1) * (2 * N - 1)) // 6)]:
        assert A.flmn_size(p.replace(storage=storage, reality=real)) == want


def test_angles_bit_exact():
    """same expression order as so3_sampling.c:224-297 gives identical bits"""
    pi = 3.141592653589793238462643383279502884197
    for L, N, ss, steer in [(8, 5, 0, 0), (8, 5, 1, 0), (9, 4, 0, 1)]:
        p = P.SO3Parameters(L=L, N=N, sampling_scheme=ss, steerable=steer)
        s = ctypes.byref(p.struct())
        for a in range(2 * L):
            want = 2.0 * a * pi / (2.0 * L - 1.0) if ss == 0 else 2.0 * a * pi / (2.0 * L)
            assert lib().so3_sampling_a2alpha(a, s) == want
            want = (2.0 * a + 1.0) * pi / (2.0 * L - 1.0) if ss == 0 else 2.0 * a * pi / (2.0 * L)
            assert lib().so3_sampling_b2beta(a, s) == want
        for g in range(2 * N):
            want = g * pi / N if steer else 2.0 * g * pi / (2.0 * N - 1.0)
            assert lib().so3_sampling_g2gamma(g, s) == want


def test_weights_closed_form():
    """w(p) = int_0^pi exp(-i p b) sin(b) db (so3_sampling.c:32-51)"""
    p = P.SO3Parameters(L=4, N=2)
    s = ctypes.byref(p.struct())
    for q in range(-9, 10):
        w = lib().so3_sampling_weight(s, q)
        if q == 1:
            assert (w.re, w.im) == (-0.0, -math.pi / 2) and math.copysign(1, w.re) < 0
        elif q == -1:
            assert (w.re, w.im) == (0.0, math.pi / 2)
        elif q % 2 == 0:
            assert (w.re, w.im) == (2.0 / (1.0 - q * q), 0.0)
        else:
            assert (w.re, w.im) == (0.0, 0.0)
        # the defining integral (Gauss-Legendre, exact to rounding)
        x, wt = np.polynomial.legendre.leggauss(60)
        b = (x + 1) * math.pi / 2
        ref = np.sum(wt * np.exp(-1j * q * b) * np.sin(b)) * math.pi / 2
        assert abs(complex(w.re, w.im) - ref) < 1e-14


def _loops(p):
    s, e, d = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    lib().so3_sampling_n_loop_values(ctypes.byref(s), ctypes.byref(e), ctypes.byref(d),
                                     ctypes.byref(p.struct()))
    return s.value, e.value, d.value


@pytest.mark.parametrize("N", [8, 7, 4, 1, 0])
def test_n_loop_values(N):
    """tests/test_convolution.c:44-132 and so3_unittest.c:36-140"""
    for real in (0, 1):
        for mode in range(5):
            p = P.SO3Parameters(L=8, N=N, reality=real, n_mode=mode)
            even = (N - 1) % 2 == 0
            if real:
                want = {0: (0, N - 1, 1), 4: (0, N - 1, 1),
                        1: (0, N - 1 if even else N - 2, 2), 2: (1, N - 1 if not even else N - 2, 2),
                        3: (N - 1, N - 1, 1)}[mode]
            else:
                want = {0: (-N + 1, N - 1, 1), 4: (-N + 1, N - 1, 1),
                        1: (-N + 1 if even else -N + 2, N - 1 if even else N - 2, 2),
                        2: (-N + 1 if not even else -N + 2, N - 1 if not even else N - 2, 2),
                        3: (-N + 1, N - 1, 2 * N - 2 if N > 1 else 1)}[mode]
            assert _loops(p) == want
            assert P.n_loop_values(p) == want


def test_is_elmn_non_zero():
    p = P.SO3Parameters(L=6, N=4, L0=1, n_mode=P.N_MODE_EVEN)
    s = ctypes.byref(p.struct())
    for n in range(-4, 5):
        for el in range(0, 7):
            for m in range(-7, 8):
                want = (n % 2 == 0 and abs(n) <= 3 and max(1, abs(n)) <= el <= 5 and abs(m) <= el)
                assert lib().so3_sampling_is_elmn_non_zero(el, m, n, s) == want
                assert lib().so3_sampling_is_elmn_non_zero_return_int(el, m, n, s) == int(want)
    assert lib().so3_sampling_is_i_in_loop_range(5, 1, 9, 2)
    assert not lib().so3_sampling_is_i_in_loop_range(4, 1, 9, 2)


def test_convolved_parameters_known_answer():
    """so3_conv_get_parameters_of_convolved_lmn on the reference's case
    (tests/test_convolution.c:191-213): f {L0=1, L=3, N=2}, g {L0=0, L=4,
    N=4} -> h {L=3, L0=1, N=3} (h.N = h.L, so3_conv.c:118-120)."""
    kat = GOLD["convolved_parameters"]
    fp = P.SO3Parameters(n_order=0, **kat["f"])
    gp = P.SO3Parameters(n_order=0, **kat["g"])
    fn = lib().so3_conv_get_parameters_of_convolved_lmn
    fn.restype = P.so3_parameters_t
    fn.argtypes = [ctypes.POINTER(P.so3_parameters_t)] * 2
    h = fn(ctypes.byref(fp.struct()), ctypes.byref(gp.struct()))
    assert (h.L, h.L0, h.N) == (kat["h"]["L"], kat["h"]["L0"], kat["h"]["N"])
    # the _void variant fills the same block
    out = P.so3_parameters_t()
    fv = lib().so3_conv_get_parameters_of_convolved_lmn_void
    fv.argtypes = [ctypes.POINTER(P.so3_parameters_t)] * 3
    fv(ctypes.byref(out), ctypes.byref(fp.struct()), ctypes.byref(gp.struct()))
    assert (out.L, out.L0, out.N) == (3, 1, 3)
