"""Python front end of the GPU library, mirroring the reference binding
src/so3/bindings.pyx:238-448 (forward, inverse, sizes, index helpers) on top
of the C ABI.  Host numpy arrays go through the drop-in entry points; torch
CUDA tensors (or raw device pointers) through the so3x_* device entry points.
"""
import ctypes

import numpy as np

from . import params as P
from ._lib import check, lib

VIA_SSHT, DIRECT = 0, 1


def _s(p):
    return ctypes.byref(p.struct())


def f_size(p):
    return lib().so3_sampling_f_size(_s(p))


def flmn_size(p):
    return lib().so3_sampling_flmn_size(_s(p))


def n_alpha(p):
    return lib().so3_sampling_nalpha(_s(p))


def n_beta(p):
    return lib().so3_sampling_nbeta(_s(p))


def n_gamma(p):
    return lib().so3_sampling_ngamma(_s(p))


def elmn2ind(el, m, n, p):
    ind = ctypes.c_int()
    fn = lib().so3_sampling_elmn2ind_real if p.reality else lib().so3_sampling_elmn2ind
    fn(ctypes.byref(ind), el, m, n, _s(p))
    return ind.value


def ind2elmn(ind, p):
    el, m, n = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    fn = lib().so3_sampling_ind2elmn_real if p.reality else lib().so3_sampling_ind2elmn
    fn(ctypes.byref(el), ctypes.byref(m), ctypes.byref(n), ind, _s(p))
    return el.value, m.value, n.value


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def inverse(flmn, p, method=VIA_SSHT):
    """flmn -> f (host arrays, synchronous)."""
    flmn = np.ascontiguousarray(flmn, dtype=np.complex128)
    if flmn.size < flmn_size(p):
        raise ValueError("flmn has %d entries, needs %d" % (flmn.size, flmn_size(p)))
    q = p if method == VIA_SSHT else p.replace(sampling_scheme=P.SO3_SAMPLING_MW, steerable=0)
    if p.reality:
        f = np.zeros(f_size(q), dtype=np.float64)
        fn = lib().so3_core_inverse_via_ssht_real if method == VIA_SSHT else lib().so3_core_inverse_direct_real
    else:
        f = np.zeros(f_size(q), dtype=np.complex128)
        fn = lib().so3_core_inverse_via_ssht if method == VIA_SSHT else lib().so3_core_inverse_direct
    fn(_ptr(f), _ptr(flmn), _s(p))
    return f


def forward(f, p, method=VIA_SSHT, flmn=None):
    """f -> flmn (host arrays, synchronous).  Blocks a via-SSHT transform
    skips (n_mode) keep the values of `flmn` when given, else zero."""
    out = np.zeros(flmn_size(p), dtype=np.complex128) if flmn is None else flmn
    if p.reality:
        f = np.ascontiguousarray(f, dtype=np.float64)
        fn = lib().so3_core_forward_via_ssht_real if method == VIA_SSHT else lib().so3_core_forward_direct_real
    else:
        f = np.ascontiguousarray(f, dtype=np.complex128)
        fn = lib().so3_core_forward_via_ssht if method == VIA_SSHT else lib().so3_core_forward_direct
    q = p if method == VIA_SSHT else p.replace(sampling_scheme=P.SO3_SAMPLING_MW, steerable=0)
    if f.size < f_size(q):
        raise ValueError("f has %d entries, needs %d" % (f.size, f_size(q)))
    fn(_ptr(out), _ptr(f), _s(p))
    return out


def s2toso3(flm, glm, p):
    """h_lmn = s2toso3_harmonic_convolution(flm, glm) for the SO(3) parameters p
    (host arrays of L^2 coefficients; reference src/so3/bindings.pyx:437-448)."""
    flm = np.ascontiguousarray(flm, dtype=np.complex128)
    glm = np.ascontiguousarray(glm, dtype=np.complex128)
    if flm.size < p.L * p.L or glm.size < p.L * p.L:
        raise ValueError("flm and glm need L^2 = %d entries" % (p.L * p.L))
    h = np.zeros(flmn_size(p), dtype=np.complex128)
    lib().so3_conv_s2toso3_harmonic_convolution(_ptr(h), _s(p), _ptr(flm), _ptr(glm))
    return h


def inverse_device(f_ptr, flmn_ptr, p, method=VIA_SSHT, stream=None):
    """Device-memory inverse on `stream` (hipStream_t as int or None)."""
    check(lib().so3x_inverse(ctypes.c_void_p(f_ptr), ctypes.c_void_p(flmn_ptr), _s(p), method,
                             ctypes.c_void_p(stream or 0)))


def forward_device(flmn_ptr, f_ptr, p, method=VIA_SSHT, stream=None):
    check(lib().so3x_forward(ctypes.c_void_p(flmn_ptr), ctypes.c_void_p(f_ptr), _s(p), method,
                             ctypes.c_void_p(stream or 0)))


def prepare(p):
    check(lib().so3x_prepare(_s(p)))


def workspace_bytes(p):
    return int(lib().so3x_workspace_bytes(_s(p)))


def device_info():
    buf = ctypes.create_string_buffer(256)
    check(lib().so3x_device_info(buf, 256))
    return buf.value.decode()


def stage_timing(on):
    check(lib().so3x_stage_timing(1 if on else 0))


def stage_times():
    """{stage: (total_ms, launches)} from the library's event profiler."""
    cap = 16
    names = (ctypes.c_char_p * cap)()
    ms = (ctypes.c_double * cap)()
    cnt = (ctypes.c_int * cap)()
    n = lib().so3x_stage_times(names, ms, cnt, cap)
    if n < 0:
        check(n)
    return {names[i].decode(): (ms[i], cnt[i]) for i in range(min(n, cap))}


def stage_reset():
    check(lib().so3x_stage_reset())


def dft_device(x_ptr, n, howmany, direction, stream=None):
    check(lib().so3x_dft(ctypes.c_void_p(x_ptr), n, howmany, direction, ctypes.c_void_p(stream or 0)))


# ---------------------------------------------------------------- batches
ITEM_INVERSE, ITEM_FORWARD, ITEM_PAIR, ITEM_S2TOSO3_INVERSE = 0, 1, 2, 3


def make_items(specs):
    """ctypes array of so3x_item_t from dicts with keys p (SO3Parameters),
    kind, method (default VIA_SSHT) and the device pointers f, flmn_in,
    flmn_out, glm (ints, 0 when unused)."""
    from ._lib import so3x_item_t
    arr = (so3x_item_t * len(specs))()
    for it, sp in zip(arr, specs):
        it.parameters = sp["p"].struct()
        it.kind = sp["kind"]
        it.method = sp.get("method", VIA_SSHT)
        it.f = sp.get("f", 0) or None
        it.flmn_in = sp.get("flmn_in", 0) or None
        it.flmn_out = sp.get("flmn_out", 0) or None
        it.glm = sp.get("glm", 0) or None
    return arr


def batch(items, lanes=0, stream=None):
    """Enqueue a batch (array from make_items) on `stream`; lanes <= 0: tuned or 4."""
    check(lib().so3x_batch(items, len(items), lanes, ctypes.c_void_p(stream or 0)))


def batch_tune(items, stream=None):
    r = lib().so3x_batch_tune(items, len(items), ctypes.c_void_p(stream or 0))
    if r < 0:
        check(r)
    return r


def graph_create(items, lanes=0):
    h = ctypes.c_longlong()
    check(lib().so3x_graph_create(items, len(items), lanes, ctypes.byref(h)))
    return h.value


def graph_launch(handle, stream=None):
    check(lib().so3x_graph_launch(handle, ctypes.c_void_p(stream or 0)))


def graph_destroy(handle):
    check(lib().so3x_graph_destroy(handle))


def ticket(stream=None):
    t = ctypes.c_longlong()
    check(lib().so3x_ticket(ctypes.c_void_p(stream or 0), ctypes.byref(t)))
    return t.value


def ticket_done(t):
    r = lib().so3x_ticket_done(t)
    if r < 0:
        check(r)
    return bool(r)


def ticket_wait(t):
    check(lib().so3x_ticket_wait(t))



def cache_limit(nbytes):
    """Bound the bytes of resident Delta tables (least recently used evicted)."""
    check(lib().so3x_cache_limit(int(nbytes)))
