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
    direct), as the reference's tests/test_so3.c:33-188 require;
  - the adjoint of the direct inverse against the conjugate transpose of
    the brute-force synthesis matrix, the adjoint of the direct forward
    against the transpose of the (quadrature-pinned) forward matrix, and the
    dot-product identities of the reference's MATLAB adjoint tests;
  - s2toso3 against the defining directional correlation
    h(rho) = int f(w) conj(g(R_rho^{-1} w)) dw on S^2, by quadrature with
    explicit rotation matrices and scipy's spherical harmonics.
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


# ------------------------------------------------------------------ adjoints
def _brute_matrix(p):
    """Columns of the synthesis map restricted to the transform's
    coefficient set, from the defining sum with explicit d-functions; a real
    signal uses the complex synthesis of the same (n >= 0) coefficients."""
    pc = p.replace(reality=0)
    l, m, n = synth.nonzero_lmn(p)
    cols = []
    for ll, mm, nn in zip(l, m, n):
        x = np.zeros(P.flmn_size(pc), complex)
        x[P.elmn2ind_vec(ll, mm, nn, pc)] = 1.0
        cols.append(brute_inverse(x, pc))
    return np.array(cols).T, (l, m, n)


ADJ_CASES = [(4, 3, 0, 0, 0, 0, 0), (5, 3, 1, 0, 0, 1, 1), (4, 3, 0, 1, 0, 0, 0), (5, 4, 0, 0, 2, 0, 0),
             (5, 4, 0, 1, 1, 1, 0), (4, 3, 0, 0, 3, 0, 1), (4, 3, 0, 0, 4, 0, 0), (4, 2, 0, 1, 2, 0, 0)]


@pytest.mark.parametrize("L,N,L0,real,mode,storage,order", ADJ_CASES)
def test_adjoint_inverse_is_brute_force_transpose(oracle, L, N, L0, real, mode, storage, order):
    """so3_adjoint_inverse_direct(f) = A^H f with A the synthesis matrix built
    from explicit d-functions (real signals: the n >= 0 half of the complex
    adjoint, the r2c transform of so3_adjoint.c:1044-1080)."""
    p = P.SO3Parameters(L=L, N=N, L0=L0, reality=real, n_mode=mode, storage=storage, n_order=order)
    A, (l, m, n) = _brute_matrix(p)
    rng = np.random.default_rng(L + 7 * N + mode)
    nf = oracle.f_size(p)
    y = rng.uniform(-1, 1, nf) if real else rng.uniform(-1, 1, nf) + 1j * rng.uniform(-1, 1, nf)
    got = oracle.adjoint_inverse(y, p)
    idx = P.elmn2ind_vec(l, m, n, p)
    assert np.abs(got[idx] - A.conj().T @ y).max() < 1e-13
    # every other stored coefficient with l >= |n| is zero (written, not left)
    rest = np.setdiff1d(np.arange(got.size), idx)
    assert np.abs(got[rest]).max(initial=0.0) == 0.0


@pytest.mark.parametrize("L,N,L0,real,mode,storage,order", ADJ_CASES)
def test_adjoint_forward_is_forward_transpose(oracle, L, N, L0, real, mode, storage, order):
    """so3_adjoint_forward_direct(x) = B^H x with B the matrix of the oracle's
    direct forward transform (pinned by the exact-quadrature test above);
    real signals: c2r semantics, Re(B^H x') with the n > 0 entries of x'
    doubled (so3_adjoint.c:1791-1825)."""
    p = P.SO3Parameters(L=L, N=N, L0=L0, reality=real, n_mode=mode, storage=storage, n_order=order)
    l, m, n = synth.nonzero_lmn(p)
    idx = P.elmn2ind_vec(l, m, n, p)
    nf = oracle.f_size(p)
    cols = []
    for j in range(nf):
        e = np.zeros(nf, float if real else complex)
        e[j] = 1.0
        cols.append(oracle.forward(e, p, "direct")[idx])
    B = np.array(cols).T
    x = synth.gen_flmn(p, 3)
    xv = x[idx]
    if real:
        ref = (B.conj().T @ (xv * np.where(n > 0, 2.0, 1.0))).real
    else:
        ref = B.conj().T @ xv
    assert np.abs(oracle.adjoint_forward(x, p) - ref).max() < 1e-13


@pytest.mark.parametrize("real", [0, 1])
def test_adjoint_dot_products_oracle(oracle, real):
    """<A x, y> = <x, A^H y> and <B y, x> = <y, B^H x> at L = N = 12 on the
    oracle's own direct pair (the reference's MATLAB checks,
    src/matlab/so3_adjoint_inverse_test.m:45-64, so3_adjoint_forward_test.m:41-65)."""
    p = P.SO3Parameters(L=12, N=12, L0=2, reality=real)
    rng = np.random.default_rng(5)
    x = synth.gen_flmn(p, 9)
    nf = oracle.f_size(p)
    y = rng.uniform(-1, 1, nf) if real else rng.uniform(-1, 1, nf) + 1j * rng.uniform(-1, 1, nf)
    l, m, n = synth.nonzero_lmn(p)
    idx = P.elmn2ind_vec(l, m, n, p)
    w = np.where(n > 0, 2.0, 1.0) if real else 1.0  # real: n > 0 coefficients count twice
    part = (lambda z: z.real) if real else (lambda z: z)  # real signals: real inner products
    lhs = np.vdot(oracle.inverse(x, p, "direct"), y)
    rhs = part(np.vdot(x[idx] * w, oracle.adjoint_inverse(y, p)[idx]))
    assert abs(lhs - rhs) < 1e-11 * abs(lhs)
    lhs = part(np.vdot(oracle.forward(y, p, "direct")[idx] * w, x[idx]))
    rhs = np.vdot(y, oracle.adjoint_forward(x, p))
    assert abs(lhs - rhs) < 1e-11 * abs(lhs)


# --------------------------------------------------------- S^2 -> SO(3) conv
def _rot_z(t):
    c, s = np.cos(t), np.sin(t)
    return np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])


def _rot_y(t):
    c, s = np.cos(t), np.sin(t)
    return np.array([[c, 0.0, s], [0.0, 1.0, 0.0], [-s, 0.0, c]])


def _s2_signal(coef, L, th, ph):
    from scipy.special import sph_harm_y
    out = np.zeros(th.shape, complex)
    for l in range(L):
        for m in range(-l, l + 1):
            if coef[l * l + l + m] != 0:
                out += coef[l * l + l + m] * sph_harm_y(l, m, th, ph)
    return out


@pytest.mark.parametrize("real", [0, 1])
def test_s2toso3_is_the_directional_correlation(oracle, real):
    """h(rho) = int_{S^2} f(w) conj(g(R_rho^{-1} w)) dw, R_rho = Rz(alpha)
    Ry(beta) Rz(gamma), evaluated by Gauss-Legendre x trapezoid quadrature
    with scipy's spherical harmonics and explicit rotation matrices, against
    the oracle's inverse of s2toso3(flm, glm) on the grid (pins the conj
    placement, the g index and the 8 pi^2/(2l+1) factor of so3_conv.c:203-208).
    g is band-limited to |m| < N, which the harmonic form assumes."""
    L, N = 6, 3
    rng = np.random.default_rng(4)
    flm = rng.uniform(-1, 1, L * L) + 1j * rng.uniform(-1, 1, L * L)
    glm = rng.uniform(-1, 1, L * L) + 1j * rng.uniform(-1, 1, L * L)
    for l in range(L):
        for m in range(-l, l + 1):
            if abs(m) >= N:
                glm[l * l + l + m] = 0
    if real:  # real signals on the sphere: f_{l,-m} = (-1)^m conj(f_lm)
        for c in (flm, glm):
            for l in range(L):
                c[l * l + l] = c[l * l + l].real
                for m in range(1, l + 1):
                    c[l * l + l - m] = (-1) ** m * np.conj(c[l * l + l + m])
    hp = P.SO3Parameters(L=L, N=N, reality=real, n_order=P.SO3_N_ORDER_ZERO_FIRST)
    h_grid = oracle.inverse(oracle.s2toso3(flm, glm, hp), hp)
    na, nb, ng = 2 * L - 1, L, 2 * N - 1
    x, w = np.polynomial.legendre.leggauss(2 * L + 2)
    nph = 4 * L
    TH, PH = np.meshgrid(np.arccos(x), 2 * np.pi * np.arange(nph) / nph, indexing="ij")
    W = np.repeat(w[:, None], nph, 1) * 2 * np.pi / nph
    v = np.stack([np.sin(TH) * np.cos(PH), np.sin(TH) * np.sin(PH), np.cos(TH)], -1)
    fv = _s2_signal(flm, L, TH, PH)
    for a, b, g in [(0, 0, 0), (3, 2, 1), (10, 5, 4), (7, 1, 2), (1, 4, 3), (9, 3, 0)]:
        R = _rot_z(2 * np.pi * a / na) @ _rot_y((2 * b + 1) * np.pi / (2 * L - 1)) @ _rot_z(2 * np.pi * g / ng)
        u = v @ R  # rows R^T v = R^{-1} v
        gv = _s2_signal(glm, L, np.arccos(np.clip(u[..., 2], -1, 1)), np.arctan2(u[..., 1], u[..., 0]))
        val = np.sum(W * fv * np.conj(gv))
        assert abs(h_grid[(g * nb + b) * na + a] - val) < 1e-13, (a, b, g)


def test_delta_tables_match_mpmath_fixtures(oracle):
    """The oracle's L = 256 Risbo and Trapani tables reproduce, through
    d^l_mn(b) = i^{n-m} sum_m' D_m'm D_m'n e^{i m' b}, the mpmath values of
    Wigner's explicit sum at the sampled beta of both grids
    (tests/golden/wigner_d_large.json, written by make_wigner_fixtures.py)."""
    import json
    import os
    fix = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "wigner_d_large.json")))
    for method in (0, 1):
        T = oracle.dl_quarter(256, method)
        for key in ("mw256", "mwss256"):
            c = fix[key]
            L = c["L"]
            for k, vals in c["d"].items():
                l, m, n = map(int, k.split(","))
                D = _full(T[l], l)
                mp = np.arange(-l, l + 1)
                for j, v in zip(c["b"], vals):
                    be = np.pi * j / L if c["ss"] else (2 * j + 1) * np.pi / (2 * L - 1)
                    d = (1j ** ((n - m) % 4)) * np.sum(D[:, m + l] * D[:, n + l] * np.exp(1j * mp * be))
                    assert abs(d - v) < 5e-14, (method, key, l, m, n, j)
    # the small-l values agree with the double-precision explicit sum
    c = fix["mw1024"]
    for j, v in zip(c["b"], c["d"]["3,3,3"]):
        assert abs(d_exact(3, 3, 3, (2 * j + 1) * np.pi / (2 * 1024 - 1)) - v) < 1e-15
