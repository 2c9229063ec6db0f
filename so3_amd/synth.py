"""Seeded synthetic Wigner coefficients.

Re and Im of every non-zero f_lmn are iid U[-1, 1) drawn from a
counter-based generator (SplitMix64 finaliser of seed and storage index), so
the same (seed, parameters) gives the same coefficients on any machine and in
any order of generation.  The set of non-zero (l, m, n) follows the
reference test generators tests/utilities.c:32-187: every n of the n-mode,
l from max(L0, |n|), |m| <= l; N_MODE_L keeps only l = |n|.  Real signals
store n >= 0 and obey f_{l,-m,0} = (-1)^m conj(f_{l,m,0}) with f_{l00} real.

The reference generator (Numerical Recipes ran2) is not reproduced: with
seed <= 0 it re-seeds on every call and returns a constant stream
(tests/utilities.c:212-224).
"""
import numpy as np

from . import params as _p

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix(x):
    x = (x + np.uint64(0x9E3779B97F4A7C15)) & _M64
    x = ((x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
    x = ((x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
    return x ^ (x >> np.uint64(31))


def uniform_pm1(seed, counters):
    """U[-1, 1) doubles for integer counters under the given seed."""
    with np.errstate(over="ignore"):
        key = _splitmix(np.uint64(seed & 0xFFFFFFFFFFFFFFFF))
        z = _splitmix(np.asarray(counters, dtype=np.uint64) ^ key)
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return 2.0 * u - 1.0


def nonzero_lmn(p):
    """(l, m, n) arrays of the coefficients that may be non-zero."""
    ls, ms, ns = [], [], []
    n_start, n_stop, n_inc = _p.n_loop_values(p)
    for n in range(n_start, n_stop + 1, n_inc):
        l_lo = max(p.L0, abs(n))
        l_hi = abs(n) if p.n_mode == _p.N_MODE_L else p.L - 1
        for el in range(l_lo, min(l_hi, p.L - 1) + 1):
            m = np.arange(-el, el + 1)
            ls.append(np.full(m.size, el)); ms.append(m); ns.append(np.full(m.size, n))
    if not ls:
        z = np.zeros(0, dtype=np.int64)
        return z, z, z
    return np.concatenate(ls), np.concatenate(ms), np.concatenate(ns)


def gen_flmn(p, seed):
    """Random flmn in the storage layout of p (complex or real signal)."""
    size = _p.flmn_size(p)
    out = np.zeros(size, dtype=np.complex128)
    l, m, n = nonzero_lmn(p)
    if l.size == 0:
        return out
    idx = _p.elmn2ind_vec(l, m, n, p)
    re = uniform_pm1(seed, 2 * idx.astype(np.uint64))
    im = uniform_pm1(seed, 2 * idx.astype(np.uint64) + np.uint64(1))
    out[idx] = re + 1j * im
    if p.reality:
        sel = n == 0
        l0, m0 = l[sel], m[sel]
        pos = m0 > 0
        ip = _p.elmn2ind_vec(l0[pos], m0[pos], np.zeros_like(l0[pos]), p)
        ineg = _p.elmn2ind_vec(l0[pos], -m0[pos], np.zeros_like(l0[pos]), p)
        sign = np.where(m0[pos] % 2 == 1, -1.0, 1.0)
        out[ineg] = sign * np.conj(out[ip])
        iz = _p.elmn2ind_vec(l0[m0 == 0], np.zeros_like(l0[m0 == 0]), np.zeros_like(l0[m0 == 0]), p)
        out[iz] = out[iz].real
    return out


def gen_flm(L, seed, real_m0=True):
    """Random S^2 coefficients f_lm (index l*l+l+m), Re/Im iid U[-1, 1)."""
    idx = np.arange(L * L, dtype=np.uint64)
    f = uniform_pm1(seed, 2 * idx) + 1j * uniform_pm1(seed, 2 * idx + np.uint64(1))
    if real_m0:
        l = np.arange(L)
        f[l * l + l] = f[l * l + l].real
    return f
