"""Port of the reference's Python tests (tests/test_pyso3.py:8-35) onto the
GPU library's binding: the round trip forward -> inverse -> forward of
arbitrary samples with the binding's defaults (negative-first, Risbo), at
the reference's tolerance (pytest.approx, relative 1e-6) and at this
build's (1e-11 absolute), and the convolution smoke test; plus
convolve_harmonic against the brute-force harmonic sum."""
import numpy as np
import pytest
from pytest import approx

import so3_amd.api as so3

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("L, N, L0", [(16, 8, 0), (16, 8, 1), (4, 1, 0), (3, 2, 0)])
@pytest.mark.parametrize("reality", [False, True])
def test_pyso3(L, N, L0, reality):
    f_p = so3.create_parameter_dict(L, N, L0, reality=reality * 1)
    f_length = so3.f_size(f_p)
    rng = np.random.default_rng(L + 10 * N + L0 + 100 * reality)
    multiplier = [1, 1] if reality else [1, 1j]
    f_before = rng.normal(size=(f_length, 2)) @ multiplier
    flmn = so3.forward(f_before, f_p)
    f_before = so3.inverse(flmn, f_p)
    flmn = so3.forward(f_before, f_p)

    f = so3.inverse(flmn, f_p)
    flmn_new = so3.forward(f, f_p)

    assert flmn_new == approx(flmn)
    assert f_before == approx(f)
    assert np.abs(flmn_new - flmn).max() < 1e-11
    assert np.abs(f_before - f).max() < 1e-11


def test_real_samples_need_the_reality_flag():
    p = so3.create_parameter_dict(8, 4)
    with pytest.raises(ValueError):
        so3.forward(np.zeros(so3.f_size(p)), p)


def test_convolve_smoke_test():
    rng = np.random.default_rng(0)
    g_p = so3.create_parameter_dict(16, 8)
    g_length = so3.f_size(g_p)
    g_before = rng.normal(size=(g_length, 2)) @ [1, 1j]
    f_before = rng.normal(size=(g_length, 2)) @ [1, 1j]
    f = so3.inverse(so3.forward(f_before, g_p), g_p)
    g = so3.inverse(so3.forward(g_before, g_p), g_p)
    h, h_p = so3.convolve(f, g_p, g, g_p)
    assert (h_p.L, h_p.N, h_p.L0) == (16, 16, 0)
    assert h.size == so3.f_size(h_p) and np.isfinite(h).all()


def test_convolve_harmonic_brute_force():
    f_p = so3.create_parameter_dict(8, 5, 1)
    g_p = so3.create_parameter_dict(9, 4)
    rng = np.random.default_rng(3)
    flmn = rng.uniform(-1, 1, so3.flmn_size(f_p)) + 1j * rng.uniform(-1, 1, so3.flmn_size(f_p))
    glmn = rng.uniform(-1, 1, so3.flmn_size(g_p)) + 1j * rng.uniform(-1, 1, so3.flmn_size(g_p))
    hlmn, h_p = so3.convolve_harmonic(flmn, f_p, glmn, g_p)
    assert hlmn.size == so3.flmn_size(h_p)
    want = np.zeros_like(hlmn)
    for n in so3.loop_over_n(h_p):
        for el in so3.loop_over_el(n, h_p):
            for m in so3.loop_over_m(el):
                acc = 0j
                for k in so3.loop_over_n(f_p):
                    if so3.is_elmn_non_zero(el, n, k, g_p):
                        acc += flmn[so3.elmn2ind(el, m, k, f_p)] * np.conj(glmn[so3.elmn2ind(el, n, k, g_p)])
                want[so3.elmn2ind(el, m, n, h_p)] = acc
    assert np.abs(hlmn - want).max() < 1e-12
