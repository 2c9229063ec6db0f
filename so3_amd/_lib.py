"""Loader and ctypes signatures of libso3_amd.so (the GPU library).

There is no CPU fallback: if the library is missing or the GPU is absent the
calls fail loudly.
"""
import ctypes
import os

from .params import so3_parameters_t

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libso3_amd.so")
_lib = None

_P = ctypes.c_void_p
_PP = ctypes.POINTER(so3_parameters_t)
_I = ctypes.c_int
_PI = ctypes.POINTER(ctypes.c_int)

# name: (restype, argtypes)
SIGNATURES = {
    # sampling (so3_sampling.h)
    "so3_sampling_f_size": (_I, [_PP]),
    "so3_sampling_n": (_I, [_PP]),
    "so3_sampling_nalpha": (_I, [_PP]),
    "so3_sampling_nbeta": (_I, [_PP]),
    "so3_sampling_ngamma": (_I, [_PP]),
    "so3_sampling_a2alpha": (ctypes.c_double, [_I, _PP]),
    "so3_sampling_b2beta": (ctypes.c_double, [_I, _PP]),
    "so3_sampling_g2gamma": (ctypes.c_double, [_I, _PP]),
    "so3_sampling_flmn_size": (_I, [_PP]),
    "so3_sampling_elmn2ind": (None, [_PI, _I, _I, _I, _PP]),
    "so3_sampling_ind2elmn": (None, [_PI, _PI, _PI, _I, _PP]),
    "so3_sampling_elmn2ind_real": (None, [_PI, _I, _I, _I, _PP]),
    "so3_sampling_ind2elmn_real": (None, [_PI, _PI, _PI, _I, _PP]),
    "so3_sampling_n_loop_values": (None, [_PI, _PI, _PI, _PP]),
    "so3_sampling_el_loop_values": (None, [_PI, _PI, _PI, _I, _PP]),
    "so3_sampling_m_loop_values": (None, [_PI, _PI, _PI, _I]),
    "so3_sampling_is_elmn_non_zero": (ctypes.c_bool, [_I, _I, _I, _PP]),
    "so3_sampling_is_elmn_non_zero_return_int": (_I, [_I, _I, _I, _PP]),
    "so3_sampling_is_i_in_loop_range": (ctypes.c_bool, [_I, _I, _I, _I]),
    # core (so3_core.h)
    "so3_core_inverse_via_ssht": (None, [_P, _P, _PP]),
    "so3_core_inverse_via_ssht_real": (None, [_P, _P, _PP]),
    "so3_core_forward_via_ssht": (None, [_P, _P, _PP]),
    "so3_core_forward_via_ssht_real": (None, [_P, _P, _PP]),
    "so3_core_inverse_direct": (None, [_P, _P, _PP]),
    "so3_core_inverse_direct_real": (None, [_P, _P, _PP]),
    "so3_core_forward_direct": (None, [_P, _P, _PP]),
    "so3_core_forward_direct_real": (None, [_P, _P, _PP]),
    # convolution (so3_conv.h)
    "so3_conv_s2toso3_harmonic_convolution": (None, [_P, _PP, _P, _P]),
    "so3_conv_harmonic_convolution": (None, [_P, _PP, _P, _PP, _P, _PP]),
    "so3_conv_convolution": (None, [_P, _PP, _P, _PP, _P, _PP]),
    "so3_conv_get_parameters_of_convolved_lmn_void": (None, [_PP, _PP, _PP]),
    # extensions (so3x.h)
    "so3x_last_error": (ctypes.c_char_p, []),
    "so3x_inverse": (_I, [_P, _P, _PP, _I, _P]),
    "so3x_forward": (_I, [_P, _P, _PP, _I, _P]),
    "so3x_inverse_noalloc": (_I, [_P, _P, _PP, _I, _P]),
    "so3x_prepare": (_I, [_PP]),
    "so3x_tune": (_I, [_PP, _I, _P, _PI, ctypes.POINTER(ctypes.c_double)]),
    "so3x_tune_get": (_I, [_PP, _I, _PI]),
    "so3x_tune_save": (_I, [ctypes.c_char_p]),
    "so3x_tune_load": (_I, [ctypes.c_char_p]),
    "so3x_tune_clear": (_I, []),
    "so3x_to_mex_layout": (_I, [_P, _P, _P, _PP, _P]),
    "so3x_from_mex_layout": (_I, [_P, _P, _P, _PP, _P]),
    "so3x_workspace_bytes": (ctypes.c_longlong, [_PP]),
    "so3x_release_all": (_I, []),
    "so3x_cache_limit": (_I, [ctypes.c_longlong]),
    "so3x_dft": (_I, [_P, _I, _I, _I, _P]),
    "so3x_device_info": (_I, [ctypes.c_char_p, _I]),
    "so3x_stage_timing": (_I, [_I]),
    "so3x_stage_times": (_I, [ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_double), _PI, _I]),
    "so3x_stage_reset": (_I, []),
}


class so3x_item_t(ctypes.Structure):
    """One item of a batch (so3x.h): parameters, kind, method, device pointers."""
    _fields_ = [("parameters", so3_parameters_t), ("kind", ctypes.c_int), ("method", ctypes.c_int),
                ("f", ctypes.c_void_p), ("flmn_in", ctypes.c_void_p), ("flmn_out", ctypes.c_void_p),
                ("glm", ctypes.c_void_p)]


_PITEM = ctypes.POINTER(so3x_item_t)
_PLL = ctypes.POINTER(ctypes.c_longlong)
SIGNATURES.update({
    "so3x_batch": (_I, [_PITEM, _I, _I, _P]),
    "so3x_batch_tune": (_I, [_PITEM, _I, _P]),
    "so3x_graph_create": (_I, [_PITEM, _I, _I, _PLL]),
    "so3x_graph_create_mode": (_I, [_PITEM, _I, _I, _I, _PLL]),
    "so3x_batch_persistent": (_I, [_PITEM, _I, _I, _P]),
    "so3x_batch_choice": (_I, [_PITEM, _I, _PI, _PI]),
    "so3x_graph_launch": (_I, [ctypes.c_longlong, _P]),
    "so3x_graph_destroy": (_I, [ctypes.c_longlong]),
    "so3x_ticket": (_I, [_P, _PLL]),
    "so3x_ticket_done": (_I, [ctypes.c_longlong]),
    "so3x_ticket_wait": (_I, [ctypes.c_longlong]),
})


class _Complex(ctypes.Structure):
    _fields_ = [("re", ctypes.c_double), ("im", ctypes.c_double)]


SIGNATURES["so3_sampling_weight"] = (_Complex, [_PP, _I])


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libso3_amd.so is not built (run: make -C so3_amd)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise RuntimeError("so3x: " + lib().so3x_last_error().decode())
    return rc
