"""ctypes front end of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and the CPU
baseline of bench.py as the checker; the library never loads it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None


class oracle_params(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in (
        "verbosity", "reality", "L0", "L", "N", "sampling_scheme", "n_order",
        "storage", "n_mode", "dl_method", "steerable")]


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = ctypes.CDLL(_LIB)
        P = ctypes.c_void_p
        for name in ("oracle_inverse_via_ssht", "oracle_inverse_via_ssht_real",
                     "oracle_forward_via_ssht", "oracle_forward_via_ssht_real",
                     "oracle_inverse_direct", "oracle_inverse_direct_real",
                     "oracle_forward_direct", "oracle_forward_direct_real",
                     "oracle_adjoint_inverse_direct", "oracle_adjoint_forward_direct",
                     "oracle_adjoint_inverse_direct_real", "oracle_adjoint_forward_direct_real"):
            getattr(_lib, name).argtypes = [P, P, ctypes.POINTER(oracle_params)]
        _lib.oracle_dl_quarter.argtypes = [ctypes.c_int, ctypes.c_int, P]
        _lib.oracle_dl_offset.argtypes = [ctypes.c_int]
        _lib.oracle_dl_offset.restype = ctypes.c_long
        _lib.oracle_dft.argtypes = [P, ctypes.c_int, ctypes.c_int]
        _lib.oracle_s2toso3.argtypes = [P, ctypes.POINTER(oracle_params), P, P]
        _lib.oracle_flmn_size.argtypes = [ctypes.POINTER(oracle_params)]
        _lib.oracle_f_size.argtypes = [ctypes.POINTER(oracle_params)]
    return _lib


def _params(p):
    q = oracle_params()
    for n, _ in oracle_params._fields_:
        setattr(q, n, int(getattr(p, n)))
    return q


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def f_size(p):
    return lib().oracle_f_size(ctypes.byref(_params(p)))


def flmn_size(p):
    return lib().oracle_flmn_size(ctypes.byref(_params(p)))


def dl_quarter(L, method):
    """list of (l+1)x(l+1) arrays D[l][m', m] = Delta^l_{m'm}(pi/2)"""
    n = lib().oracle_dl_offset(L)
    out = np.zeros(n + 1)
    lib().oracle_dl_quarter(L, method, _ptr(out))
    tabs = []
    for l in range(L):
        o = lib().oracle_dl_offset(l)
        tabs.append(out[o:o + (l + 1) ** 2].reshape(l + 1, l + 1))
    return tabs


def dft(x, sign):
    y = np.ascontiguousarray(x, dtype=np.complex128).copy()
    lib().oracle_dft(_ptr(y), y.size, sign)
    return y


def _call(name, out, inp, p):
    q = _params(p)
    getattr(lib(), name)(_ptr(out), _ptr(inp), ctypes.byref(q))
    return out


def inverse(flmn, p, method="via_ssht"):
    flmn = np.ascontiguousarray(flmn, dtype=np.complex128)
    if p.reality:
        f = np.zeros(f_size(p), dtype=np.float64)
        return _call("oracle_inverse_%s_real" % method, f, flmn, p)
    f = np.zeros(f_size(p), dtype=np.complex128)
    return _call("oracle_inverse_%s" % method, f, flmn, p)


def forward(f, p, method="via_ssht", out=None):
    flmn = np.zeros(flmn_size(p), dtype=np.complex128) if out is None else out
    if p.reality:
        return _call("oracle_forward_%s_real" % method, flmn, np.ascontiguousarray(f, dtype=np.float64), p)
    return _call("oracle_forward_%s" % method, flmn, np.ascontiguousarray(f, dtype=np.complex128), p)


def adjoint_inverse(f, p):
    flmn = np.zeros(flmn_size(p), dtype=np.complex128)
    if p.reality:
        return _call("oracle_adjoint_inverse_direct_real", flmn, np.ascontiguousarray(f, dtype=np.float64), p)
    return _call("oracle_adjoint_inverse_direct", flmn, np.ascontiguousarray(f, dtype=np.complex128), p)


def adjoint_forward(flmn, p):
    flmn = np.ascontiguousarray(flmn, dtype=np.complex128)
    if p.reality:
        f = np.zeros(f_size(p), dtype=np.float64)
        return _call("oracle_adjoint_forward_direct_real", f, flmn, p)
    f = np.zeros(f_size(p), dtype=np.complex128)
    return _call("oracle_adjoint_forward_direct", f, flmn, p)


def s2toso3(flm, glm, hp):
    h = np.zeros(flmn_size(hp), dtype=np.complex128)
    flm = np.ascontiguousarray(flm, dtype=np.complex128)
    glm = np.ascontiguousarray(glm, dtype=np.complex128)
    lib().oracle_s2toso3(_ptr(h), ctypes.byref(_params(hp)), _ptr(flm), _ptr(glm))
    return h


def time_sample(p, ns, gamma_every=1):
    """CPU seconds of the sampled via-SSHT work: (per-n array, dl build,
    gamma DFT over all columns, extrapolated from every gamma_every-th)."""
    ns = np.ascontiguousarray(ns, dtype=np.int32)
    t = np.zeros(ns.size)
    out = np.zeros(2)
    q = _params(p)
    lib().oracle_time_sample.argtypes = [ctypes.POINTER(oracle_params), ctypes.c_void_p, ctypes.c_int,
                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    lib().oracle_time_sample(ctypes.byref(q), _ptr(ns), ns.size, _ptr(t), _ptr(out), int(gamma_every))
    return t, out[0], out[1]
