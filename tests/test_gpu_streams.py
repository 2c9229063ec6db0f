"""Concurrency and lifetime of the device-memory entry points.

* so3x_inverse / so3x_forward on two non-blocking streams with different
  inputs, enqueued back to back: both results equal the oracle (the single
  transforms share one workspace, ordered between streams by events).
* A captured batch graph owns its workspace and pins its Delta tables: after
  a larger eager batch (which regrows the shared lane workspaces) and a
  table build under a zero cache limit (which evicts every unpinned table),
  its replays are still bit-identical to the first one.
"""
import numpy as np
import pytest
import torch

import so3_amd.api as A
from so3_amd import params as P
from so3_amd import synth

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


def test_two_streams_share_the_workspace_safely(oracle):
    p1 = P.SO3Parameters(L=40, N=24, n_order=P.SO3_N_ORDER_ZERO_FIRST)
    p2 = P.SO3Parameters(L=48, N=12, n_order=P.SO3_N_ORDER_ZERO_FIRST)
    x1, x2 = synth.gen_flmn(p1, 1), synth.gen_flmn(p2, 2)
    r1, r2 = oracle.inverse(x1, p1), oracle.inverse(x2, p2)
    g2 = oracle.inverse(synth.gen_flmn(p2, 3), p2)
    rb2 = oracle.forward(g2, p2)
    d1, d2 = torch.from_numpy(x1).to(DEV), torch.from_numpy(x2).to(DEV)
    dg2 = torch.from_numpy(g2).to(DEV)
    f1 = torch.zeros(A.f_size(p1), dtype=torch.complex128, device=DEV)
    f2 = torch.zeros(A.f_size(p2), dtype=torch.complex128, device=DEV)
    b2 = torch.zeros(A.flmn_size(p2), dtype=torch.complex128, device=DEV)
    A.prepare(p1)
    A.prepare(p2)
    torch.cuda.synchronize()
    s1, s2 = torch.cuda.Stream(device=DEV), torch.cuda.Stream(device=DEV)
    for _ in range(3):
        A.inverse_device(f1.data_ptr(), d1.data_ptr(), p1, A.VIA_SSHT, s1.cuda_stream)
        A.inverse_device(f2.data_ptr(), d2.data_ptr(), p2, A.VIA_SSHT, s2.cuda_stream)
        A.forward_device(b2.data_ptr(), dg2.data_ptr(), p2, A.VIA_SSHT, s2.cuda_stream)
        A.inverse_device(f1.data_ptr(), d1.data_ptr(), p1, A.VIA_SSHT, s1.cuda_stream)
    torch.cuda.synchronize()
    for got, ref in ((f1, r1), (f2, r2), (b2, rb2)):
        g = got.cpu().numpy()
        assert np.abs(g - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())


def _items(specs, seed):
    out, keep = [], []
    for i, (L, N, dl) in enumerate(specs):
        p = P.SO3Parameters(L=L, N=N, dl_method=dl, n_order=P.SO3_N_ORDER_ZERO_FIRST)
        flmn = torch.from_numpy(synth.gen_flmn(p, seed + i)).to(DEV)
        f = torch.zeros(A.f_size(p), dtype=torch.complex128, device=DEV)
        back = torch.zeros_like(flmn)
        out.append({"p": p, "kind": A.ITEM_PAIR, "f": f.data_ptr(), "flmn_in": flmn.data_ptr(),
                    "flmn_out": back.data_ptr()})
        keep.append((flmn, f, back))
    return A.make_items(out), keep


def test_graph_survives_workspace_growth_and_eviction():
    from so3_amd._lib import lib
    sid = torch.cuda.current_stream().cuda_stream
    # start from an empty table cache, so that the L = 130 build below is the
    # largest table and evicts the graph's tables unless they are pinned
    assert lib().so3x_release_all() == 0
    items, keep = _items([(20, 8, 0), (33, 5, 1), (17, 17, 0)], 50)
    g = A.graph_create(items, lanes=2)
    try:
        A.graph_launch(g, sid)
        torch.cuda.synchronize()
        first = [(f.clone(), b.clone()) for _, f, b in keep]
        # a larger eager batch regrows the shared lane workspaces
        big, keep_big = _items([(96, 48, 0), (80, 80, 1), (64, 30, 0)], 60)
        A.batch(big, lanes=2, stream=sid)
        torch.cuda.synchronize()
        # a table build under a zero limit evicts every table not pinned
        # (L = 130 is larger than every cached table, so it is built)
        A.cache_limit(0)
        A.prepare(P.SO3Parameters(L=130, N=2))
        for rep in range(3):
            for _, f, b in keep:
                f.zero_()
                b.zero_()
            A.graph_launch(g, sid)
            torch.cuda.synchronize()
            for (f0, b0), (_, f, b) in zip(first, keep):
                assert torch.equal(f, f0) and torch.equal(b, b0), rep
    finally:
        A.cache_limit(32 << 30)
        A.graph_destroy(g)


def test_graph_replays_on_two_streams_are_ordered():
    """two replays of one graph on different streams run one after the other
    (they share the graph's workspace): both leave the same outputs"""
    items, keep = _items([(24, 12, 0), (30, 6, 0)], 70)
    s1, s2 = torch.cuda.Stream(device=DEV), torch.cuda.Stream(device=DEV)
    g = A.graph_create(items, lanes=2)
    try:
        A.graph_launch(g, s1.cuda_stream)
        torch.cuda.synchronize()
        first = [(f.clone(), b.clone()) for _, f, b in keep]
        A.graph_launch(g, s1.cuda_stream)
        A.graph_launch(g, s2.cuda_stream)
        torch.cuda.synchronize()
        for (f0, b0), (_, f, b) in zip(first, keep):
            assert torch.equal(f, f0) and torch.equal(b, b0)
    finally:
        A.graph_destroy(g)


@pytest.mark.parametrize("L,N", [(40, 12), (200, 64)])
def test_host_pointer_paths_agree(L, N):
    """The drop-in entry points on pageable host arrays (staged through
    pinned buffers in 64 MiB chunks: L = 200, N = 64 moves 162 MB of samples)
    and on pinned host arrays (direct copies) give the same bits as the
    device-pointer call."""
    p = P.SO3Parameters(L=L, N=N, n_order=P.SO3_N_ORDER_ZERO_FIRST)
    flmn = synth.gen_flmn(p, 9)
    f_pageable = A.inverse(flmn, p)
    pin_in = torch.empty(flmn.size, dtype=torch.complex128, pin_memory=True)
    pin_in.numpy()[:] = flmn
    pin_out = torch.empty(A.f_size(p), dtype=torch.complex128, pin_memory=True)
    import ctypes
    from so3_amd._lib import lib
    lib().so3_core_inverse_via_ssht(ctypes.c_void_p(pin_out.data_ptr()), ctypes.c_void_p(pin_in.data_ptr()),
                                    ctypes.byref(p.struct()))
    d_in = torch.from_numpy(flmn).to(DEV)
    d_out = torch.zeros(A.f_size(p), dtype=torch.complex128, device=DEV)
    A.inverse_device(d_out.data_ptr(), d_in.data_ptr(), p)
    torch.cuda.synchronize()
    ref = d_out.cpu().numpy()
    assert np.array_equal(f_pageable, ref)
    assert np.array_equal(pin_out.numpy(), ref)
    back = A.forward(f_pageable, p)
    assert np.abs(back - flmn).max() < 1e-11
