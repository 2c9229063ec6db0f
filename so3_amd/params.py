"""Parameter block and enums of the SO(3) transform API.

Mirrors so3_parameters_t (reference include/so3/so3_types.h:130-201) as a
ctypes structure and the Python-side SO3Parameters object of the reference
binding (src/so3/bindings.pyx:96-130), including its defaults
(NEGATIVE_FIRST order, Risbo recursion).  The vectorised index helpers give
the same integers as so3_sampling_elmn2ind / _real (so3_sampling.c:369-597).
"""
import ctypes

import numpy as np

SO3_N_ORDER_ZERO_FIRST, SO3_N_ORDER_NEGATIVE_FIRST = 0, 1
SO3_STORAGE_PADDED, SO3_STORAGE_COMPACT = 0, 1
N_MODE_ALL, N_MODE_EVEN, N_MODE_ODD, N_MODE_MAXIMUM, N_MODE_L = 0, 1, 2, 3, 4
SO3_N_MODE_ALL, SO3_N_MODE_EVEN, SO3_N_MODE_ODD, SO3_N_MODE_MAXIMUM, SO3_N_MODE_L = 0, 1, 2, 3, 4
SO3_SAMPLING_MW, SO3_SAMPLING_MW_SS = 0, 1
SSHT_DL_RISBO, SSHT_DL_TRAPANI = 0, 1


class so3_parameters_t(ctypes.Structure):
    _fields_ = [(name, ctypes.c_int) for name in (
        "verbosity", "reality", "L0", "L", "N", "sampling_scheme", "n_order",
        "storage", "n_mode", "dl_method", "steerable")]


class SO3Parameters:
    """Transform parameters with the reference binding's defaults."""

    def __init__(self, L=32, N=4, L0=0, verbosity=0, reality=0,
                 sampling_scheme=SO3_SAMPLING_MW, n_order=SO3_N_ORDER_NEGATIVE_FIRST,
                 storage=SO3_STORAGE_PADDED, n_mode=N_MODE_ALL,
                 dl_method=SSHT_DL_RISBO, steerable=0):
        self.L, self.N, self.L0 = int(L), int(N), int(L0)
        self.verbosity, self.reality = int(verbosity), int(reality)
        self.sampling_scheme, self.n_order = int(sampling_scheme), int(n_order)
        self.storage, self.n_mode = int(storage), int(n_mode)
        self.dl_method, self.steerable = int(dl_method), int(steerable)

    def from_dict(self, d):
        """Fill the fields from a dict (or a so3_parameters_t) with the same keys
        (bindings.pyx:118-130)."""
        get = (lambda k: d[k]) if isinstance(d, dict) else (lambda k: getattr(d, k))
        for name, _ in so3_parameters_t._fields_:
            setattr(self, name, int(get(name)))
        return self

    def struct(self):
        s = so3_parameters_t()
        for name, _ in so3_parameters_t._fields_:
            setattr(s, name, getattr(self, name))
        return s

    def replace(self, **kw):
        q = SO3Parameters(**{k: getattr(self, k) for k, _ in so3_parameters_t._fields_})
        for k, v in kw.items():
            setattr(q, k, int(v))
        return q

    def __repr__(self):
        return "SO3Parameters(%s)" % ", ".join(
            "%s=%d" % (k, getattr(self, k)) for k, _ in so3_parameters_t._fields_)


def nalpha(p):
    return 2 * p.L - 1 if p.sampling_scheme == SO3_SAMPLING_MW else 2 * p.L


def nbeta(p):
    return p.L if p.sampling_scheme == SO3_SAMPLING_MW else p.L + 1


def ngamma(p):
    return p.N if p.steerable else 2 * p.N - 1


def f_size(p):
    return nalpha(p) * nbeta(p) * ngamma(p)


def flmn_size(p):
    L, N = p.L, p.N
    if p.storage == SO3_STORAGE_PADDED:
        return N * L * L if p.reality else (2 * N - 1) * L * L
    if p.reality:
        return N * (6 * L * L - (N - 1) * (2 * N - 1)) // 6
    return (2 * N - 1) * (3 * L * L - N * (N - 1)) // 3


def n_loop_values(p):
    N = p.N
    top_even = (N - 1) % 2 == 0
    if p.reality:
        return {N_MODE_ALL: (0, N - 1, 1), N_MODE_L: (0, N - 1, 1),
                N_MODE_EVEN: (0, N - 1 if top_even else N - 2, 2),
                N_MODE_ODD: (1, N - 1 if not top_even else N - 2, 2),
                N_MODE_MAXIMUM: (N - 1, N - 1, 1)}[p.n_mode]
    return {N_MODE_ALL: (-N + 1, N - 1, 1), N_MODE_L: (-N + 1, N - 1, 1),
            N_MODE_EVEN: ((-N + 1, N - 1, 2) if top_even else (-N + 2, N - 2, 2)),
            N_MODE_ODD: ((-N + 1, N - 1, 2) if not top_even else (-N + 2, N - 2, 2)),
            N_MODE_MAXIMUM: (-N + 1, N - 1, 2 * N - 2 if N > 1 else 1)}[p.n_mode]


def _block_base(n, order, p):
    """index(l, m, n) = base(n) + l*l + l + m in both storages."""
    L2, N = p.L * p.L, p.N
    n = np.asarray(n, dtype=np.int64)
    an = np.abs(n)
    if p.storage == SO3_STORAGE_PADDED:
        if order == SO3_N_ORDER_ZERO_FIRST:
            return np.where(n < 0, -2 * n - 1, 2 * n) * L2
        return (N - 1 + n) * L2
    if order == SO3_N_ORDER_ZERO_FIRST:
        off = (2 * an - 1) * (3 * L2 - an * (an - 1)) // 3
        off = off + np.where(n >= 0, L2 - n * n, 0)
    else:
        off = (N - 1 + n) * L2 - (2 * N - 1) * (N - 1) * N // 6
        off = off + np.where(n <= 0, an * (2 * an + 1) * (an + 1) // 6,
                             -(an * (2 * an - 1) * (an - 1) // 6))
    return off - n * n


def elmn2ind_vec(l, m, n, p):
    """Storage index of (l, m, n); real signals use the _real map."""
    l = np.asarray(l, dtype=np.int64)
    m = np.asarray(m, dtype=np.int64)
    if p.reality:
        base = _block_base(n, SO3_N_ORDER_NEGATIVE_FIRST, p) - _block_base(0, SO3_N_ORDER_NEGATIVE_FIRST, p)
    else:
        base = _block_base(n, p.n_order, p)
    return base + l * l + l + m
