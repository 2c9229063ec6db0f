"""Batched execution (so3x_batch, graphs, tuning, tickets) against the plain
per-transform calls: every item of a mixed batch must produce exactly the
same bits as the same transform run alone, eagerly, through a captured and
replayed graph, and at any lane count."""
import numpy as np
import pytest
import torch

import so3_amd.api as A
from so3_amd import params as P
from so3_amd import synth

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")

# (L, N, sampling, real, dl, kind, method)
SPECS = [
    (12, 5, 0, 0, 0, A.ITEM_PAIR, A.VIA_SSHT),
    (9, 9, 1, 1, 0, A.ITEM_PAIR, A.VIA_SSHT),
    (16, 3, 0, 0, 1, A.ITEM_INVERSE, A.VIA_SSHT),
    (7, 1, 0, 1, 0, A.ITEM_FORWARD, A.VIA_SSHT),
    (20, 8, 0, 0, 0, A.ITEM_PAIR, A.DIRECT),
    (33, 6, 0, 0, 0, A.ITEM_S2TOSO3_INVERSE, A.VIA_SSHT),
    (5, 4, 1, 0, 0, A.ITEM_PAIR, A.VIA_SSHT),
    (24, 24, 0, 1, 0, A.ITEM_PAIR, A.VIA_SSHT),
]


def _setup():
    items, refs, outs = [], [], []
    rng = np.random.default_rng(3)
    for i, (L, N, ss, real, dl, kind, meth) in enumerate(SPECS):
        p = P.SO3Parameters(L=L, N=N, sampling_scheme=ss, reality=real, dl_method=dl,
                            n_order=P.SO3_N_ORDER_ZERO_FIRST)
        if meth == A.DIRECT:
            p = P.SO3Parameters(L=L, N=N, sampling_scheme=0, reality=real, dl_method=dl,
                                n_order=P.SO3_N_ORDER_ZERO_FIRST)
        fdt = torch.float64 if real else torch.complex128
        f = torch.zeros(A.f_size(p), dtype=fdt, device=DEV)
        fo = torch.zeros(A.flmn_size(p), dtype=torch.complex128, device=DEV)
        sp = {"p": p, "kind": kind, "method": meth, "f": f.data_ptr(), "flmn_out": fo.data_ptr()}
        keep = [f, fo]
        if kind == A.ITEM_S2TOSO3_INVERSE:
            flm = rng.uniform(-1, 1, L * L) + 1j * rng.uniform(-1, 1, L * L)
            glm = rng.uniform(-1, 1, L * L) + 1j * rng.uniform(-1, 1, L * L)
            dflm = torch.from_numpy(flm).to(DEV)
            dglm = torch.from_numpy(glm).to(DEV)
            sp["flmn_in"], sp["glm"] = dflm.data_ptr(), dglm.data_ptr()
            keep += [dflm, dglm]
            h = A.s2toso3(flm, glm, p)
            ref = {"f": torch.from_numpy(A.inverse(h, p)).to(DEV), "flmn_out": torch.from_numpy(h).to(DEV)}
        elif kind == A.ITEM_FORWARD:
            fin = A.inverse(synth.gen_flmn(p, 20 + i), p, meth)
            f.copy_(torch.from_numpy(fin).to(DEV))
            ref = {"flmn_out": torch.from_numpy(A.forward(fin, p, meth)).to(DEV)}
        else:
            flmn = synth.gen_flmn(p, 20 + i)
            din = torch.from_numpy(flmn).to(DEV)
            sp["flmn_in"] = din.data_ptr()
            keep.append(din)
            fr = A.inverse(flmn, p, meth)
            ref = {"f": torch.from_numpy(fr).to(DEV)}
            if kind == A.ITEM_PAIR:
                ref["flmn_out"] = torch.from_numpy(A.forward(fr, p, meth)).to(DEV)
        items.append(sp)
        refs.append(ref)
        outs.append({"f": f, "flmn_out": fo, "keep": keep, "kind": kind})
    return A.make_items(items), refs, outs


def _clear(outs):
    for o in outs:
        if o["kind"] != A.ITEM_FORWARD:  # a forward item reads f
            o["f"].zero_()
        o["flmn_out"].zero_()


def _check(refs, outs):
    for r, o in zip(refs, outs):
        for k, v in r.items():
            assert torch.equal(o[k], v), (o["kind"], k, (o[k] - v).abs().max().item())


@pytest.mark.parametrize("lanes", [1, 3, 8, 0])
def test_batch_matches_alone(lanes):
    items, refs, outs = _setup()
    _clear(outs)
    A.batch(items, lanes=lanes, stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    _check(refs, outs)


def test_graph_replays_match_eager():
    items, refs, outs = _setup()
    g = A.graph_create(items, lanes=4)
    try:
        s = torch.cuda.current_stream()
        for rep in range(3):
            _clear(outs)
            torch.cuda.synchronize()
            A.graph_launch(g, s.cuda_stream)
            # eager work of the program between replays
            x = torch.ones(1 << 20, dtype=torch.float64, device=DEV) * rep
            torch.cuda.synchronize()
            assert float(x.sum()) == rep * (1 << 20)
            _check(refs, outs)
    finally:
        A.graph_destroy(g)


def test_tune_and_tickets():
    items, refs, outs = _setup()
    s = torch.cuda.current_stream().cuda_stream
    best = A.batch_tune(items, stream=s)
    assert best in (1, 2, 4, 8)
    _clear(outs)
    A.batch(items, lanes=0, stream=s)
    t = A.ticket(s)
    A.ticket_wait(t)
    _check(refs, outs)
    t2 = A.ticket(s)
    torch.cuda.synchronize()
    assert A.ticket_done(t2)
    A.ticket_wait(t2)
