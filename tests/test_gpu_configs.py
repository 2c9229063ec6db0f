"""Every benchmark configuration at its own size, with absolute checks.

* L = N = 256 (MW and MW_SS, complex and real, Risbo and Trapani Delta):
  the GPU inverse of a sparse coefficient set at 300 grid points against
  the defining sum f = sum (2l+1)/(8 pi^2) f_lmn e^{i m a} d^l_mn(b) e^{i n g}
  with d^l_mn(b) from high-precision fixtures (tests/golden/wigner_d_large.json,
  Wigner's explicit sum in mpmath), then the round trip; and full random
  transforms with Risbo and with Trapani tables agreeing.
* L = 1024: S^2 -> SO(3) convolution then the inverse at N = 8 (the plain
  entry points and the batch item), and the standalone N = 4 transform,
  against the same fixture-based sum.
* The 96-item uneven batch of the benchmark and a 40-request slice of the
  serving stream: every item bit-identical to the same transform alone.

Tolerances: 1e-12 relative to max |f| (the GPU-vs-reference bar), round
trips 1e-11; batched and served results bit-identical.
"""
import json
import os

import numpy as np
import pytest
import torch

import so3_amd.api as A
from so3_amd import params as P
from so3_amd import serve, synth

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = json.load(open(os.path.join(HERE, "golden", "wigner_d_large.json")))
DEV = torch.device("cuda:0")

COEFS = [(255, 255, -255), (255, -3, 254), (200, -17, 33), (128, 0, 0), (97, 96, -50),
         (1, 1, -1), (254, -200, 0), (60, 5, 60)]
REAL_COEFS = [(255, 255, 255), (255, -3, 254), (200, -17, 33), (128, 0, 0), (97, 96, 50),
              (1, 1, 1), (254, -200, 7), (60, 5, 60)]


def _grid(L, N, ss, steer=False):
    na, nb = (2 * L, L + 1) if ss else (2 * L - 1, L)
    return na, nb, 2 * N - 1


def _check_sparse(f, vals, fix, L, N, ss, real, rng, npts=300):
    """f (GPU grid) against the defining sum at npts points whose beta
    index is one of the fixture's sampled b; returns (max err, max |ref|)."""
    na, nb, ng = _grid(L, N, ss)
    assert f.size == na * nb * ng
    bs = fix["b"]
    err, fmax = 0.0, 0.0
    for _ in range(npts):
        jb = int(rng.integers(0, len(bs)))
        b = bs[jb]
        a, g = int(rng.integers(0, na)), int(rng.integers(0, ng))
        al, ga = 2 * np.pi * a / na, 2 * np.pi * g / ng
        ref = 0j
        for (l, m, n), c in vals.items():
            d = fix["d"]["%d,%d,%d" % (l, m, n)][jb]
            t = c * (2 * l + 1) / (8 * np.pi ** 2) * np.exp(1j * m * al) * d * np.exp(1j * n * ga)
            ref += t + np.conj(t) if (real and n > 0) else t
        if real:
            ref = ref.real
        got = f[(g * nb + b) * na + a]
        err = max(err, abs(got - ref))
        fmax = max(fmax, abs(ref))
    return err, fmax


@pytest.mark.parametrize("dl", [P.SSHT_DL_RISBO, P.SSHT_DL_TRAPANI])
@pytest.mark.parametrize("ss,real", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_c1_sparse_absolute(dl, ss, real):
    """L = N = 256 synthesis against the mpmath d-functions, both Delta methods"""
    L = N = 256
    p = P.SO3Parameters(L=L, N=N, reality=real, sampling_scheme=ss, dl_method=dl,
                        n_order=P.SO3_N_ORDER_ZERO_FIRST)
    rng = np.random.default_rng(31 + dl + 2 * ss + 4 * real)
    flmn = np.zeros(P.flmn_size(p), dtype=np.complex128)
    vals = {}
    for (l, m, n) in (REAL_COEFS if real else COEFS):
        v = rng.uniform(-1, 1) + 1j * rng.uniform(-1, 1)
        if real and n == 0 and m == 0:
            v = v.real + 0j
        vals[(l, m, n)] = v
        flmn[P.elmn2ind_vec(np.array([l]), np.array([m]), np.array([n]), p)[0]] = v
    f = A.inverse(flmn, p)
    err, fmax = _check_sparse(f, vals, FIX["mwss256" if ss else "mw256"], L, N, ss, real, rng)
    assert err <= 1e-12 * max(fmax, 1.0), (err, fmax)
    back = A.forward(f, p)
    assert np.abs(back - flmn).max() < 1e-11


@pytest.mark.parametrize("ss,real", [(0, 0), (0, 1), (1, 0)])
def test_c1_risbo_equals_trapani(ss, real):
    """Full random L = N = 256 transforms with the Risbo and the Trapani
    tables agree (catches a table that is wrong but orthogonal, which a
    round trip with the same table cannot)"""
    p = P.SO3Parameters(L=256, N=256, reality=real, sampling_scheme=ss, n_order=P.SO3_N_ORDER_ZERO_FIRST)
    q = p.replace(dl_method=P.SSHT_DL_TRAPANI)
    flmn = synth.gen_flmn(p, 1)
    fr = A.inverse(flmn, p)
    ft = A.inverse(flmn, q)
    assert np.abs(fr - ft).max() <= 1e-12 * np.abs(fr).max()
    br = A.forward(fr, p)
    bt = A.forward(fr, q)
    assert np.abs(br - bt).max() <= 1e-12 * np.abs(br).max()


def _conv_inputs():
    L = 1024
    rng = np.random.default_rng(2)
    flm = np.zeros(L * L, complex)
    glm = np.zeros(L * L, complex)
    for (l, m) in FIX["flm_1024"]:
        flm[l * l + l + m] = rng.uniform(-1, 1) + 1j * rng.uniform(-1, 1)
    for (l, n) in FIX["glm_1024_n8"]:
        glm[l * l + l + n] = rng.uniform(-1, 1) + 1j * rng.uniform(-1, 1)
    vals = {}
    for (l, m) in FIX["flm_1024"]:
        for (lg, n) in FIX["glm_1024_n8"]:
            if lg == l:
                vals[(l, m, n)] = flm[l * l + l + m] * np.conj(glm[l * l + l + n]) * 8 * np.pi ** 2 / (2 * l + 1)
    return flm, glm, vals


def test_c2_s2toso3_inverse_absolute():
    """h = s2toso3(flm, glm) at L = 1024, N = 8, then the inverse to the
    2047 x 1024 x 15 grid, against the defining sum with mpmath d-functions"""
    hp = P.SO3Parameters(L=1024, N=8, n_order=P.SO3_N_ORDER_ZERO_FIRST)
    flm, glm, vals = _conv_inputs()
    h = A.s2toso3(flm, glm, hp)
    # h holds exactly the products of the sparse inputs
    nz = np.flatnonzero(h)
    assert len(nz) == len(vals)
    for (l, m, n), v in vals.items():
        assert abs(h[A.elmn2ind(l, m, n, hp)] - v) <= 1e-15 * abs(v)
    f = A.inverse(h, hp)
    rng = np.random.default_rng(3)
    err, fmax = _check_sparse(f, vals, FIX["mw1024"], 1024, 8, 0, 0, rng)
    assert err <= 1e-12 * max(fmax, 1.0), (err, fmax)
    # the same through the batch item (s2toso3 + inverse in one item)
    f_d = torch.zeros(A.f_size(hp), dtype=torch.complex128, device=DEV)
    h_d = torch.zeros(A.flmn_size(hp), dtype=torch.complex128, device=DEV)
    flm_d = torch.from_numpy(flm).to(DEV)
    glm_d = torch.from_numpy(glm).to(DEV)
    items = A.make_items([{"p": hp, "kind": A.ITEM_S2TOSO3_INVERSE, "f": f_d.data_ptr(),
                           "flmn_in": flm_d.data_ptr(), "glm": glm_d.data_ptr(), "flmn_out": h_d.data_ptr()}])
    A.batch(items, lanes=1, stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert np.array_equal(h_d.cpu().numpy(), h)
    assert np.array_equal(f_d.cpu().numpy(), f)


def test_c2b_l1024_n4_absolute():
    """standalone complex L = 1024, N = 4 transform: sparse synthesis against
    the mpmath d-functions, then the round trip"""
    p = P.SO3Parameters(L=1024, N=4, n_order=P.SO3_N_ORDER_ZERO_FIRST)
    rng = np.random.default_rng(4)
    flmn = np.zeros(P.flmn_size(p), complex)
    vals = {}
    for (l, m, n) in FIX["coefs_1024_n4"]:
        v = rng.uniform(-1, 1) + 1j * rng.uniform(-1, 1)
        vals[(l, m, n)] = v
        flmn[A.elmn2ind(l, m, n, p)] = v
    f = A.inverse(flmn, p)
    err, fmax = _check_sparse(f, vals, FIX["mw1024"], 1024, 4, 0, 0, rng)
    assert err <= 1e-12 * max(fmax, 1.0), (err, fmax)
    back = A.forward(f, p)
    assert np.abs(back - flmn).max() < 1e-11


def _batched_workload(seed=4, count=96):
    """The benchmark's uneven batch: L uniform on {32, 48, ..., 384}, N
    uniform on [1, min(L, 64)], MW / MW_SS and complex / real 50/50."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        L = int(rng.choice(np.arange(32, 385, 16)))
        N = int(rng.integers(1, min(L, 64) + 1))
        ss = int(rng.integers(0, 2))
        real = int(rng.integers(0, 2))
        out.append(P.SO3Parameters(L=L, N=N, sampling_scheme=ss, reality=real, n_order=P.SO3_N_ORDER_ZERO_FIRST))
    return out


def test_c3_full_batch_bit_identical():
    """all 96 items of the uneven batch: batched (eager and graph replay, in
    the lanes and the persistent mode) == alone, bit for bit"""
    ps = _batched_workload()
    stream = torch.cuda.current_stream()
    sid = stream.cuda_stream
    specs, outs, refs, keep = [], [], [], []
    for i, p in enumerate(ps):
        flmn = torch.from_numpy(synth.gen_flmn(p, 100 + i)).to(DEV)
        fdt = torch.float64 if p.reality else torch.complex128
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
        outs.append((f, back))
        keep.append(flmn)
    items = A.make_items(specs)
    A.batch(items, lanes=0, stream=sid)
    torch.cuda.synchronize()
    for i, ((f, b), (rf, rb)) in enumerate(zip(outs, refs)):
        assert torch.equal(f, rf) and torch.equal(b, rb), (i, ps[i])
    for mode in (A.BATCH_LANES, A.BATCH_PERSISTENT):
        g = A.graph_create(items, lanes=8, mode=mode)
        try:
            for f, b in outs:
                f.zero_()
                b.zero_()
            torch.cuda.synchronize()
            A.graph_launch(g, sid)
            torch.cuda.synchronize()
            for i, ((f, b), (rf, rb)) in enumerate(zip(outs, refs)):
                assert torch.equal(f, rf) and torch.equal(b, rb), (mode, i, ps[i])
        finally:
            A.graph_destroy(g)
    # the persistent mode, eager
    for f, b in outs:
        f.zero_()
        b.zero_()
    A.batch_persistent(items, lanes=8, stream=sid)
    torch.cuda.synchronize()
    for i, ((f, b), (rf, rb)) in enumerate(zip(outs, refs)):
        assert torch.equal(f, rf) and torch.equal(b, rb), ("persistent", i, ps[i])


def test_c4_serving_slice_bit_identical():
    """40 requests of the serving stream (L up to 1024, s2toso3 + inverse
    among them), served with continuous batching: each equals the same
    request run alone through the plain entry points"""
    reqs = serve.request_mix(40, 400.0)
    assert max(r.p.L for r in reqs) > 512
    fn = serve.random_inputs(5)
    srv = serve.Server(hbm_budget=48 << 30, lanes=4, max_batch=16, streams=3)
    srv.prepare(reqs)
    srv.run(reqs, fn, keep_outputs=True)
    for r in reqs:
        d = fn(r)
        if r.kind == A.ITEM_PAIR:
            f = A.inverse(d["flmn"], r.p)
            ref = {"f": f, "flmn_out": A.forward(f, r.p)}
        else:
            h = A.s2toso3(d["flm"], d["glm"], r.p)
            ref = {"f": A.inverse(h, r.p), "flmn_out": h}
        for k, v in ref.items():
            assert np.array_equal(r.outputs[k], v), (r.rid, r.p, k)
