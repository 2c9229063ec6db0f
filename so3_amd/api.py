"""Python front end of the GPU library, mirroring the reference binding
src/so3/bindings.pyx:96-448 (parameters, sizes, index and loop helpers,
forward, inverse, convolutions) on top of the C ABI, with the binding's
defaults (negative-first n order, Risbo).  Host numpy arrays go through the
drop-in entry points; torch CUDA tensors (or raw device pointers) through the
so3x_* device entry points.

One deliberate difference: convolve_harmonic runs the harmonic-space
convolution (so3_conv_harmonic_convolution) on coefficients.  The reference
binding calls the real-space so3_conv_convolution there (bindings.pyx:427-434),
which would read coefficients as samples and write f_size(h) samples into a
buffer of flmn_size(h) entries.
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


_ENUMS = {
    "sampling_scheme": {"SO3_SAMPLING_MW": P.SO3_SAMPLING_MW, "SO3_SAMPLING_MW_SS": P.SO3_SAMPLING_MW_SS},
    "n_order": {"SO3_N_ORDER_NEGATIVE_FIRST": P.SO3_N_ORDER_NEGATIVE_FIRST,
                "SO3_N_ORDER_ZERO_FIRST": P.SO3_N_ORDER_ZERO_FIRST},
    "storage": {"SO3_STORAGE_PADDED": P.SO3_STORAGE_PADDED, "SO3_STORAGE_COMPACT": P.SO3_STORAGE_COMPACT},
    "n_mode": {"SO3_N_MODE_ALL": P.N_MODE_ALL, "SO3_N_MODE_EVEN": P.N_MODE_EVEN, "SO3_N_MODE_ODD": P.N_MODE_ODD,
               "SO3_N_MODE_MAXIMUM": P.N_MODE_MAXIMUM, "SO3_N_MODE_L": P.N_MODE_L},
    "dl_method": {"SSHT_DL_RISBO": P.SSHT_DL_RISBO, "SSHT_DL_TRAPANI": P.SSHT_DL_TRAPANI},
}
# a string that names no value selects the binding's fallback (its else branch)
_FALLBACK = {"sampling_scheme": P.SO3_SAMPLING_MW_SS, "n_order": P.SO3_N_ORDER_ZERO_FIRST,
             "storage": P.SO3_STORAGE_COMPACT, "n_mode": P.N_MODE_L, "dl_method": P.SSHT_DL_TRAPANI}


def create_parameter_dict(L, N, L0=0, verbosity=0, reality=0, sampling_scheme_str="SO3_SAMPLING_MW",
                          n_order_str="SO3_N_ORDER_NEGATIVE_FIRST", storage_str="SO3_STORAGE_PADDED",
                          n_mode_str="SO3_N_MODE_ALL", dl_method_str="SSHT_DL_RISBO", steerable=0):
    """SO3Parameters from enum names (bindings.pyx:136-217)."""
    def pick(field, s):
        return _ENUMS[field].get(s, _FALLBACK[field])
    return P.SO3Parameters(L=L, N=N, L0=L0, verbosity=verbosity, reality=reality,
                           sampling_scheme=pick("sampling_scheme", sampling_scheme_str),
                           n_order=pick("n_order", n_order_str), storage=pick("storage", storage_str),
                           n_mode=pick("n_mode", n_mode_str), dl_method=pick("dl_method", dl_method_str),
                           steerable=steerable)


def get_loop_n_values(p):
    """(n_start, n_stop, n_inc) of so3_sampling_n_loop_values"""
    a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    lib().so3_sampling_n_loop_values(ctypes.byref(a), ctypes.byref(b), ctypes.byref(c), _s(p))
    return a.value, b.value, c.value


def loop_over_n(p):
    n_start, n_stop, n_inc = get_loop_n_values(p)
    yield from range(n_start, n_stop + 1, n_inc)


def get_loop_el_values(n, p):
    """(el_start, el_stop, el_inc) of so3_sampling_el_loop_values"""
    a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    lib().so3_sampling_el_loop_values(ctypes.byref(a), ctypes.byref(b), ctypes.byref(c), n, _s(p))
    return a.value, b.value, c.value


def loop_over_el(n, p):
    el_start, el_stop, el_inc = get_loop_el_values(n, p)
    yield from range(el_start, el_stop + 1, el_inc)


def get_loop_m_values(el):
    """(m_start, m_stop, m_inc) of so3_sampling_m_loop_values"""
    a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    lib().so3_sampling_m_loop_values(ctypes.byref(a), ctypes.byref(b), ctypes.byref(c), el)
    return a.value, b.value, c.value


def loop_over_m(el):
    m_start, m_stop, m_inc = get_loop_m_values(el)
    yield from range(m_start, m_stop + 1, m_inc)


def is_elmn_non_zero(el, m, n, p):
    return bool(lib().so3_sampling_is_elmn_non_zero_return_int(el, m, n, _s(p)))


def inverse(flmn, p, method=VIA_SSHT):
    """flmn -> f (host arrays, synchronous); a real signal when p.reality."""
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
    f = np.asarray(f)
    if not np.iscomplexobj(f) and not p.reality:
        raise ValueError("f has a real data type but reality flag not set")
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


def s2toso3_harmonic_convolution(h_parameters, flm, glm):
    """The reference binding's name and argument order (bindings.pyx:437-448)."""
    return s2toso3(flm, glm, h_parameters)


def get_convolved_parameters(f_parameters, g_parameters):
    """Parameters of h = f (*) g (so3_conv_get_parameters_of_convolved_lmn_void;
    bindings.pyx:368-380): L = min, L0 = max, N = L."""
    h = P.so3_parameters_t()
    lib().so3_conv_get_parameters_of_convolved_lmn_void(ctypes.byref(h), _s(f_parameters), _s(g_parameters))
    return P.SO3Parameters().from_dict(h)


def get_convolved_parameter_dict(f_parameters_dict, g_parameters_dict):
    """Deprecated dict form (bindings.pyx:361-366)."""
    print("This function is depreaciated use SO3Parameters class to define parameters instead")
    return get_convolved_parameters(P.SO3Parameters().from_dict(f_parameters_dict),
                                    P.SO3Parameters().from_dict(g_parameters_dict))


def convolve(f, f_parameters, g, g_parameters):
    """h = f (*) g on the sampling grids (so3_conv_convolution: direct forward
    of f and g, harmonic product, direct inverse); returns (h, h_parameters)
    (bindings.pyx:382-408).  f, g and h are complex signals."""
    h_parameters = get_convolved_parameters(f_parameters, g_parameters)
    f = np.ascontiguousarray(f, dtype=np.complex128)
    g = np.ascontiguousarray(g, dtype=np.complex128)
    for x, q, name in ((f, f_parameters, "f"), (g, g_parameters, "g")):
        need = f_size(q.replace(reality=0, sampling_scheme=P.SO3_SAMPLING_MW, steerable=0))
        if x.size < need:
            raise ValueError("%s has %d samples, needs %d" % (name, x.size, need))
    h = np.zeros(f_size(h_parameters.replace(reality=0, sampling_scheme=P.SO3_SAMPLING_MW, steerable=0)),
                 dtype=np.complex128)
    lib().so3_conv_convolution(_ptr(h), _s(h_parameters), _ptr(f), _s(f_parameters), _ptr(g), _s(g_parameters))
    return h, h_parameters


def convolve_harmonic(flmn, f_parameters, glmn, g_parameters):
    """hlmn = flmn (*) glmn in harmonic space (so3_conv_harmonic_convolution,
    so3_conv.c:44-90); returns (hlmn, h_parameters)."""
    h_parameters = get_convolved_parameters(f_parameters, g_parameters)
    flmn = np.ascontiguousarray(flmn, dtype=np.complex128)
    glmn = np.ascontiguousarray(glmn, dtype=np.complex128)
    if flmn.size < flmn_size(f_parameters) or glmn.size < flmn_size(g_parameters):
        raise ValueError("flmn / glmn shorter than their flmn_size")
    hlmn = np.zeros(flmn_size(h_parameters), dtype=np.complex128)
    lib().so3_conv_harmonic_convolution(_ptr(hlmn), _s(h_parameters), _ptr(flmn), _s(f_parameters),
                                        _ptr(glmn), _s(g_parameters))
    return hlmn, h_parameters


def inverse_device(f_ptr, flmn_ptr, p, method=VIA_SSHT, stream=None):
    """Device-memory inverse on `stream` (hipStream_t as int or None)."""
    check(lib().so3x_inverse(ctypes.c_void_p(f_ptr), ctypes.c_void_p(flmn_ptr), _s(p), method,
                             ctypes.c_void_p(stream or 0)))


def forward_device(flmn_ptr, f_ptr, p, method=VIA_SSHT, stream=None):
    check(lib().so3x_forward(ctypes.c_void_p(flmn_ptr), ctypes.c_void_p(f_ptr), _s(p), method,
                             ctypes.c_void_p(stream or 0)))


def to_mex_layout_device(re_ptr, im_ptr, f_ptr, p, stream=None):
    """f (library layout) -> MATLAB layout re / im on the device (im 0 for real p)."""
    check(lib().so3x_to_mex_layout(ctypes.c_void_p(re_ptr), ctypes.c_void_p(im_ptr or 0), ctypes.c_void_p(f_ptr),
                                   _s(p), ctypes.c_void_p(stream or 0)))


def from_mex_layout_device(f_ptr, re_ptr, im_ptr, p, stream=None):
    """MATLAB layout re / im -> f (library layout) on the device (im 0: zero imaginary parts)."""
    check(lib().so3x_from_mex_layout(ctypes.c_void_p(f_ptr), ctypes.c_void_p(re_ptr), ctypes.c_void_p(im_ptr or 0),
                                     _s(p), ctypes.c_void_p(stream or 0)))


def prepare(p):
    check(lib().so3x_prepare(_s(p)))


def tune(p, method=VIA_SSHT, stream=None):
    """Run the (L, N)-keyed autotuner for p; returns ({"corder_inv", "corder_fwd",
    "pw", "gamma_engine"}, {"contract_inv_ms", "contract_fwd_ms", "beta_fwd_ms",
    "gamma_ms"}).  gamma_engine: 1 workgroup, 2 256-point register FFT, 0 where
    the gamma length 2N-1 is outside [65, 127] (gamma_ms is then 0)."""
    ch = (ctypes.c_int * 4)()
    ms = (ctypes.c_double * 4)()
    check(lib().so3x_tune(_s(p), method, ctypes.c_void_p(stream or 0), ch, ms))
    return ({"corder_inv": ch[0], "corder_fwd": ch[1], "pw": ch[2], "gamma_engine": ch[3]},
            {"contract_inv_ms": ms[0], "contract_fwd_ms": ms[1], "beta_fwd_ms": ms[2], "gamma_ms": ms[3]})


def tune_get(p, method=VIA_SSHT):
    ch = (ctypes.c_int * 4)()
    check(lib().so3x_tune_get(_s(p), method, ch))
    return {"corder_inv": ch[0], "corder_fwd": ch[1], "pw": ch[2], "gamma_engine": ch[3]}


def tune_save(path):
    check(lib().so3x_tune_save(str(path).encode()))


def tune_clear():
    check(lib().so3x_tune_clear())


def tune_load(path):
    r = lib().so3x_tune_load(str(path).encode())
    if r < 0:
        check(r)
    return r


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


BATCH_AUTO, BATCH_LANES, BATCH_PERSISTENT = 0, 1, 2


def batch_persistent(items, lanes=0, stream=None):
    """Enqueue a batch in the persistent mode (grouped, queue-fed contractions)."""
    check(lib().so3x_batch_persistent(items, len(items), lanes, ctypes.c_void_p(stream or 0)))


def batch_choice(items):
    """(mode, lanes) in force for this batch's shape."""
    m, l = ctypes.c_int(), ctypes.c_int()
    check(lib().so3x_batch_choice(items, len(items), ctypes.byref(m), ctypes.byref(l)))
    return m.value, l.value


def batch_tune(items, stream=None):
    r = lib().so3x_batch_tune(items, len(items), ctypes.c_void_p(stream or 0))
    if r < 0:
        check(r)
    return r


def graph_create(items, lanes=0, mode=BATCH_AUTO):
    h = ctypes.c_longlong()
    check(lib().so3x_graph_create_mode(items, len(items), lanes, mode, ctypes.byref(h)))
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
