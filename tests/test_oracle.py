"""Pins the CPU oracle before it is trusted as the GPU checker.

Each check is chosen so that a plausible mistake (a dropped term, a wrong
sign or index, a transposed operand) fails it:
  - DFTs against numpy.fft (pocketfft) at the lengths the transforms use;
  - Delta^l(pi/2) against exact values of Wigner's explicit sum (rational
    arithmetic), orthogonality, Risbo vs Trapani;
  - the inverse transform against brute-force synthesis
    f = sum (2l+1)/(8 pi^2) f_lmn exp(i m a) d^l_mn(b) exp(i n g)
    with exact d-functions (pins the absolute convention);
  - the forward transform of arbitrary (not band-limited) samples against
    the exact integral of their MW / MW_SS trigonometric interpolant
    (pins extension sign, quadrature weights and the half-sample phase);
  - round trips and the agreement of the two formulations (via SSHT vs
    direct), as the reference's tests/test_so3.c:33-188 require.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

import so3_amd.params as P
from so3_amd import synth

pytestmark = pytest.mark.filterwarnings("ignore")


@pytest.mark.parametrize("n", [1, 2, 3, 5, 7, 8, 15, 31, 61, 73, 127, 255, 511, 1021, 1024, 2047])
def test_dft_vs_numpy(oracle, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ref = np.fft.fft(x)
    scale = np.abs(ref).max()
    assert np.abs(oracle.dft(x, -1) - ref).max() / scale < 4e-15 * max(1, math.log2(n))
    assert np.abs(oracle.dft(x, +1) - np.fft.ifft(x) * n).max() / scale < 4e-15 * max(1, math.log2(n))


def _fact(k):
    return math.factorial(k)


def d_exact(l, mp, m, beta=None):
    """Wigner d^l_{mp m}(beta) from the explicit sum; at beta = pi/2 the
    cos/sin powers are exactly 2^-l, so the result is an exact rational
    times sqrt of an integer, evaluated in Fractions then rounded once."""
    if beta is None:
        s = Fraction(0)
        for k in range(0, 2 * l + 1):
            a, b, c, e = l + m - k, k, l - k - mp, k - m + mp
            if min(a, b, c, e) < 0:
                continue
            s += Fraction((-1) ** (k - m + mp), _fact(a) * _fact(b) * _fact(c) * _fact(e))
        rad = _fact(l + m) * _fact(l - m) * _fact(l + mp) * _fact(l - mp)
        return float(s * Fraction(math.isqrt(rad)) / 2 ** l) if math.isqrt(rad) ** 2 == rad else \
            float(s / 2 ** l) * math.sqrt(rad)
    cb, sb = math.cos(beta / 2), math.sin(beta / 2)
    s = 0.0
    for k in range(0, 2 * l + 1):
        a, b, c, e = l + m - k, k, l - k - mp, k - m + mp
        if min(a, b, c, e) < 0:
            continue
        s += (-1) ** (k - m + mp) * cb ** (2 * l - 2 * k + m - mp) * sb ** (2 * k - m + mp) / \
            (_fact(a) * _fact(b) * _fact(c) * _fact(e))
    return s * math.sqrt(_fact(l + m) * _fact(l - m) * _fact(l + mp) * _fact(l - mp))


@pytest.mark.parametrize("method", [0, 1])
def test_delta_exact_small_l(oracle, method):
    T = oracle.dl_quarter(20, method)
    for l in range(20):
        for mp in range(l + 1):
            for m in range(l + 1):
                assert abs(T[l][mp, m] - d_exact(l, mp, m)) < 1e-14, (l, mp, m)


def _full(Tl, l):
    """full (2l+1)^2 Delta^l(pi/2) from the quarter via the symmetries
    D_{m',-m} = (-1)^{l+m'} D_{m',m}, D_{-m',m} = (-1)^{l+m} D_{m',m}"""
    D = np.zeros((2 * l + 1, 2 * l + 1))
    for mp in range(-l, l + 1):
        for m in range(-l, l + 1):
            s = 1.0
            a, b = mp, m
            if a < 0:
                a = -a
                s *= (-1) ** (l + b)
            if b < 0:
                b = -b
                s *= (-1) ** (l + a)
            D[mp + l, m + l] = s * Tl[a, b]
    return D


def test_delta_orthogonal_and_methods_agree(oracle):
    L = 256
    Tr = oracle.dl_quarter(L, 0)
    Tt = oracle.dl_quarter(L, 1)
    assert max(np.abs(a - b).max() for a, b in zip(Tr, Tt)) < 1e-13
    for l in (1, 7, 64, 255):
        D = _full(Tr[l], l)
        assert np.abs(D.T @ D - np.eye(2 * l + 1)).max() < 1e-13
        # Delta^l_{00} = P_l(0)
        p0 = 0.0 if l % 2 else (-1) ** (l // 2) * math.comb(l, l // 2) / 2 ** l
        assert abs(D[l, l] - p0) < 1e-14


def brute_inverse(flmn, p):
    """f on the sampling grid by the defining sum with explicit d-functions"""
    L, N = p.L, p.N
    ss = p.sampling_scheme
    na, nb = (2 * L - 1, L) if ss == 0 else (2 * L, L + 1)
    ng = 2 * N - 1
    al = 2 * np.pi * np.arange(na) / na
    be = (2 * np.arange(nb) + 1) * np.pi / (2 * L - 1) if ss == 0 else np.pi * np.arange(nb) / L
    ga = 2 * np.pi * np.arange(ng) / ng
    f = np.zeros((ng, nb, na), complex)
    ns = range(0, N) if p.reality else range(-(N - 1), N)
    for n in ns:
        for l in range(max(p.L0, abs(n)), L):
            for m in range(-l, l + 1):
                c = flmn[P.elmn2ind_vec(l, m, n, p)] * (2 * l + 1) / (8 * np.pi ** 2)
                d = np.array([d_exact(l, m, n, b) for b in be])
                term = c * np.exp(1j * m * al)[None, None, :] * d[None, :, None] * \
                    np.exp(1j * n * ga)[:, None, None]
                if p.reality and n > 0:
                    # real signal: the n < 0 half is the conjugate partner
                    f += term + np.conj(term)
                else:
                    f += term
    return f.ravel().real if p.reality else f.ravel()


@pytest.mark.parametrize("ss", [0, 1])
@pytest.mark.parametrize("real", [0, 1])
@pytest.mark.parametrize("L,N,L0", [(4, 3, 0), (5, 5, 1)])
def test_inverse_brute_force(oracle, ss, real, L, N, L0):
    p = P.SO3Parameters(L=L, N=N, L0=L0, sampling_scheme=ss, reality=real, n_order=0)
    flmn = synth.gen_flmn(p, 11)
    f = oracle.inverse(flmn, p)
    ref = brute_inverse(flmn, p)
    assert np.abs(f - ref).max() < 1e-13


def exact_forward(f, p):
    """f_lmn = int f~ D^l_mn sin(b) da db dg, f~ the trigonometric
    interpolant of the samples: degree L-1 in alpha, N-1 in gamma, and in
    beta the interpolant of the samples extended to [0, 2 pi) by
    F(2pi - b) = (-1)^{m+n} F(b) (MW: 2L-1 points, MW_SS: 2L points with the
    Nyquist term dropped).  The beta integral is done by Gauss-Legendre."""
    L, N = p.L, p.N
    ss = p.sampling_scheme
    na, nb = (2 * L - 1, L) if ss == 0 else (2 * L, L + 1)
    ng = 2 * N - 1
    f = f.reshape(ng, nb, na)
    Fab = np.fft.fft(np.fft.fft(f, axis=2), axis=0) / (na * ng)  # [n slot][b][m slot]
    x, w = np.polynomial.legendre.leggauss(8 * L + 40)
    bq = (x + 1) * np.pi / 2
    wq = w * np.pi / 2
    out = {}
    ns = range(0, N) if p.reality else range(-(N - 1), N)
    for n in ns:
        for m in range(-(L - 1), L):
            col = Fab[n % ng, :, m % na]
            s = (-1) ** ((m + n) % 2)
            if ss == 0:
                ext = np.concatenate([col, s * col[L - 2::-1]])
                th = (2 * np.arange(2 * L - 1) + 1) * np.pi / (2 * L - 1)
            else:
                ext = np.concatenate([col, s * col[L - 1:0:-1]])
                th = 2 * np.pi * np.arange(2 * L) / (2 * L)
            ks = np.arange(-(L - 1), L)
            c = np.array([np.mean(ext * np.exp(-1j * k * th)) for k in ks])
            Pb = (c[None, :] * np.exp(1j * ks[None, :] * bq[:, None])).sum(1)
            for l in range(max(p.L0, abs(m), abs(n)), L):
                d = np.array([d_exact(l, m, n, b) for b in bq])
                out[(l, m, n)] = 4 * np.pi ** 2 * np.sum(wq * Pb * d * np.sin(bq))
    return out


@pytest.mark.parametrize("ss", [0, 1])
@pytest.mark.parametrize("real", [0, 1])
@pytest.mark.parametrize("method", ["via_ssht", "direct"])
def test_forward_is_exact_quadrature(oracle, ss, real, method):
    if method == "direct" and ss:
        pytest.skip("direct transforms are MW only")
    L, N = 4, 3
    p = P.SO3Parameters(L=L, N=N, sampling_scheme=ss, reality=real, n_order=0)
    rng = np.random.default_rng(3)
    size = oracle.f_size(p)
    f = rng.uniform(-1, 1, size) if real else rng.uniform(-1, 1, size) + 1j * rng.uniform(-1, 1, size)
    flmn = oracle.forward(f, p, method)
    ref = exact_forward(f, p)
    for (l, m, n), v in ref.items():
        assert abs(flmn[P.elmn2ind_vec(l, m, n, p)] - v) < 1e-13, (l, m, n)


CONFIGS = [(16, 16, 0, 0, 0, 0), (16, 5, 2, 1, 0, 1), (12, 12, 0, 0, 1, 0), (9, 4, 1, 1, 1, 1)]


@pytest.mark.parametrize("L,N,L0,ss,real,dl", CONFIGS)
def test_round_trip(oracle, L, N, L0, ss, real, dl):
    p = P.SO3Parameters(L=L, N=N, L0=L0, sampling_scheme=ss, reality=real, dl_method=dl)
    flmn = synth.gen_flmn(p, 5)
    back = oracle.forward(oracle.inverse(flmn, p), p)
    assert np.abs(back - flmn).max() < 1e-13


@pytest.mark.parametrize("real", [0, 1])
@pytest.mark.parametrize("storage,order,mode", [(0, 0, 0), (1, 1, 1), (0, 1, 2), (1, 0, 3)])
def test_direct_equals_via_ssht(oracle, real, storage, order, mode):
    """tests/test_so3.c:66-96,135-188, on the whole f and on arbitrary f"""
    p = P.SO3Parameters(L=8, N=8, reality=real, storage=storage, n_order=order, n_mode=mode,
                        dl_method=1)
    flmn = synth.gen_flmn(p, 10)
    f1 = oracle.inverse(flmn, p, "via_ssht")
    f2 = oracle.inverse(flmn, p, "direct")
    assert np.abs(f1 - f2).max() < 1e-13
    rng = np.random.default_rng(1)
    g = rng.uniform(-1, 1, f1.shape) + (0 if real else 1j * rng.uniform(-1, 1, f1.shape))
    z = np.zeros(oracle.flmn_size(p), complex)
    b1 = oracle.forward(g, p, "via_ssht", out=z.copy())
    b2 = oracle.forward(g, p, "direct", out=z.copy())
    # the two agree on every coefficient of an active block (direct zeroes the others)
    l, m, n = synth.nonzero_lmn(p)
    idx = P.elmn2ind_vec(l, m, n, p)
    assert np.abs(b1[idx] - b2[idx]).max() < 1e-13


def test_real_signal_is_real(oracle):
    """a real-signal coefficient set synthesises the real part of the
    complex synthesis of its conjugate-completed set, whose imaginary part
    vanishes (reality condition f_{l,-m,-n} = (-1)^{m+n} conj f_lmn)"""
    pr = P.SO3Parameters(L=7, N=4, reality=1)
    pc = pr.replace(reality=0, n_order=P.SO3_N_ORDER_NEGATIVE_FIRST)
    fr = synth.gen_flmn(pr, 2)
    fc = np.zeros(P.flmn_size(pc), complex)
    l, m, n = synth.nonzero_lmn(pr)
    v = fr[P.elmn2ind_vec(l, m, n, pr)]
    fc[P.elmn2ind_vec(l, m, n, pc)] = v
    sgn = np.where((m + n) % 2 == 1, -1.0, 1.0)
    fc[P.elmn2ind_vec(l, -m, -n, pc)] = sgn * np.conj(v)
    a = oracle.inverse(fr, pr)
    b = oracle.inverse(fc, pc)
    assert np.abs(b.imag).max() < 1e-13
    assert np.abs(a - b.real).max() < 1e-13
