"""so3x_inverse_noalloc: the allocation-free inverse on the library's static
workspace against so3x_inverse (the same kernels on the allocated
workspace), bit for bit, on two calls in a row (the second smaller), with
the device's free memory unchanged by the calls; and its errors for a
transform too large for the static workspace and for unprepared parameters.
"""
import ctypes

import numpy as np
import pytest
import torch

import so3_amd.api as A
from so3_amd import params as P
from so3_amd import synth
from so3_amd._lib import lib

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


def _noalloc(f, flmn, p, sid, method=A.VIA_SSHT):
    return lib().so3x_inverse_noalloc(ctypes.c_void_p(f.data_ptr()), ctypes.c_void_p(flmn.data_ptr()),
                                      ctypes.byref(p.struct()), method, ctypes.c_void_p(sid))


def _case(p, seed):
    flmn = torch.from_numpy(synth.gen_flmn(p, seed)).to(DEV)
    fdt = torch.float64 if p.reality else torch.complex128
    ref = torch.zeros(A.f_size(p), dtype=fdt, device=DEV)
    out = torch.full((A.f_size(p),), float("nan"), dtype=fdt, device=DEV)
    return flmn, ref, out


def test_noalloc_matches_allocating_inverse():
    sid = torch.cuda.current_stream().cuda_stream
    z = P.SO3_N_ORDER_ZERO_FIRST
    cases = [P.SO3Parameters(L=64, N=48, n_order=z),                       # first call
             P.SO3Parameters(L=24, N=9, sampling_scheme=1, n_order=z),     # second, smaller
             P.SO3Parameters(L=40, N=40, reality=1, dl_method=1, n_order=z)]
    bufs = []
    for k, p in enumerate(cases):
        A.prepare(p)
        flmn, ref, out = _case(p, 7 + k)
        A.inverse_device(ref.data_ptr(), flmn.data_ptr(), p, A.VIA_SSHT, sid)
        bufs.append((p, flmn, ref, out))
    torch.cuda.synchronize()
    for k, (p, flmn, ref, out) in enumerate(bufs):
        assert _noalloc(out, flmn, p, sid) == 0, lib().so3x_last_error().decode()
        torch.cuda.synchronize()
        assert torch.equal(out, ref), (k, p)
    # repeated calls allocate nothing: the device's free memory is the same
    # before and after them (read with no other work in between)
    for _, _, _, out in bufs:
        out.fill_(float("nan"))
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    for p, flmn, ref, out in bufs:
        assert _noalloc(out, flmn, p, sid) == 0
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info()[0] == free0
    for p, flmn, ref, out in bufs:
        assert torch.equal(out, ref)


def test_noalloc_errors():
    sid = torch.cuda.current_stream().cuda_stream
    big = P.SO3Parameters(L=128, N=128, n_order=P.SO3_N_ORDER_ZERO_FIRST)
    A.prepare(big)
    flmn, _, out = _case(big, 1)
    assert _noalloc(out, flmn, big, sid) != 0
    assert "static workspace" in lib().so3x_last_error().decode()
    fresh = P.SO3Parameters(L=29, N=7, L0=3, n_order=P.SO3_N_ORDER_ZERO_FIRST)
    flmn, _, out = _case(fresh, 2)
    assert _noalloc(out, flmn, fresh, sid) != 0
    assert "prepare" in lib().so3x_last_error().decode()
    torch.cuda.synchronize()
    assert torch.isnan(out.real if out.is_complex() else out).all()  # nothing was written
