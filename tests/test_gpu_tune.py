"""The (L, N)-keyed autotuner on the GPU.

* The contraction workgroup order never changes a result: with order 1
  forced for both contractions (through a loaded table) the transforms are
  bit-identical to order 0, at shapes with one and many lane blocks.
* so3x_tune picks its choices; the tuned transform still matches the oracle
  (another weight-convolution length changes only the rounding), and a
  batch run after tuning is bit-identical to the tuned single transforms.
"""
import numpy as np
import pytest
import torch

import so3_amd.api as A
from so3_amd import params as P
from so3_amd import synth

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")
Z = P.SO3_N_ORDER_ZERO_FIRST


def _pair(p, flmn):
    f = A.inverse(flmn, p)
    return f, A.forward(f, p)


@pytest.mark.parametrize("L,N,real,ss", [(64, 33, 0, 0), (150, 12, 1, 1), (256, 256, 0, 0), (300, 5, 0, 1),
                                         (96, 96, 1, 0)])
def test_contraction_order_is_bit_identical(tmp_path, L, N, real, ss):
    p = P.SO3Parameters(L=L, N=N, reality=real, sampling_scheme=ss, n_order=Z)
    flmn = synth.gen_flmn(p, 3)
    t = tmp_path / "t.txt"
    t.write_text("%d %d %d %d 0 0 0\n" % (L, N, real, ss))
    A.tune_load(t)
    f0, b0 = _pair(p, flmn)
    t.write_text("%d %d %d %d 1 1 0\n" % (L, N, real, ss))
    A.tune_load(t)
    assert A.tune_get(p)["corder_inv"] == 1 and A.tune_get(p)["corder_fwd"] == 1
    f1, b1 = _pair(p, flmn)
    assert np.array_equal(f0, f1) and np.array_equal(b0, b1)
    t.write_text("%d %d %d %d 0 0 0\n" % (L, N, real, ss))
    A.tune_load(t)


@pytest.mark.parametrize("L,N,real,ss", [(40, 20, 0, 0), (120, 8, 1, 1), (200, 30, 0, 1), (90, 40, 1, 0),
                                         (120, 45, 0, 0)])
def test_tuned_transform_matches_oracle(oracle, L, N, real, ss):
    p = P.SO3Parameters(L=L, N=N, reality=real, sampling_scheme=ss, n_order=Z)
    choice, ms = A.tune(p)
    assert choice["corder_inv"] in (0, 1) and choice["corder_fwd"] in (0, 1)
    assert 4 * L - 3 <= choice["pw"] <= 4096
    # the gamma engine is a choice for gamma lengths 2N-1 in [65, 127] only
    # (shorter ones take the direct DFT)
    assert choice["gamma_engine"] in ((1, 2) if 65 <= 2 * N - 1 <= 127 else (0,))
    assert (ms["gamma_ms"] > 0) == (choice["gamma_engine"] != 0)
    assert all(v > 0 for k, v in ms.items() if k != "gamma_ms")
    assert A.tune_get(p) == choice
    flmn = synth.gen_flmn(p, 8)
    f = A.inverse(flmn, p)
    ref = oracle.inverse(flmn, p)
    assert np.abs(f - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())
    rng = np.random.default_rng(L)
    g = rng.uniform(-1, 1, f.size) + (0 if real else 1j * rng.uniform(-1, 1, f.size))
    got = A.forward(g, p)
    want = oracle.forward(g, p)
    assert np.abs(got - want).max() <= 1e-12 * max(1.0, np.abs(want).max())


@pytest.fixture(autouse=True)
def _fresh_table():
    A.tune_clear()
    yield
    A.tune_clear()


def test_batch_after_tuning_matches_alone(tmp_path):
    shapes = [(48, 16, 0, 0), (130, 7, 1, 0), (64, 64, 0, 1)]
    specs, refs, keep = [], [], []
    sid = torch.cuda.current_stream().cuda_stream
    for i, (L, N, real, ss) in enumerate(shapes):
        p = P.SO3Parameters(L=L, N=N, reality=real, sampling_scheme=ss, n_order=Z)
        A.tune(p)
        flmn = torch.from_numpy(synth.gen_flmn(p, 20 + i)).to(DEV)
        fdt = torch.float64 if real else torch.complex128
        f = torch.zeros(A.f_size(p), dtype=fdt, device=DEV)
        back = torch.zeros_like(flmn)
        A.inverse_device(f.data_ptr(), flmn.data_ptr(), p, A.VIA_SSHT, sid)
        A.forward_device(back.data_ptr(), f.data_ptr(), p, A.VIA_SSHT, sid)
        torch.cuda.synchronize()
        refs.append((f.clone(), back.clone()))
        f.zero_()
        back.zero_()
        specs.append({"p": p, "kind": A.ITEM_PAIR, "f": f.data_ptr(), "flmn_in": flmn.data_ptr(),
                      "flmn_out": back.data_ptr()})
        keep.append((flmn, f, back))
    A.batch(A.make_items(specs), lanes=3, stream=sid)
    torch.cuda.synchronize()
    for (_, f, b), (rf, rb) in zip(keep, refs):
        assert torch.equal(f, rf) and torch.equal(b, rb)
    path = tmp_path / "tuned.txt"
    A.tune_save(path)
    assert A.tune_load(path) >= len(shapes)
