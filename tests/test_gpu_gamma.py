"""The gamma DFT at every short length class against the CPU oracle.

Gamma lengths 2N-1 <= 63 run as a direct DFT (gamma_short.hip: one column
per thread, or per pair of lanes for complex lengths above 31); lengths in
[65, 127] on the workgroup engine (smooth lengths, directly) or the
256-point register FFT (Bluestein), either one forced through the tuning
table in the second test; longer ones on the other FFT engines.  Each case runs the inverse of seeded coefficients and the
forward of arbitrary samples (which pins the forward DFT on data that is not
band-limited) through the C ABI, complex and real, MW and MW_SS, against the
oracle at 1e-12 relative to max|input|.
"""
import numpy as np
import pytest

import so3_amd.api as A
import so3_amd.params as P
from so3_amd import synth

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(1.0, np.abs(b).max())


def _check(oracle, N, real, ss):
    L = N + 3
    p = P.SO3Parameters(L=L, N=N, sampling_scheme=ss, reality=real, n_order=P.SO3_N_ORDER_ZERO_FIRST)
    flmn = synth.gen_flmn(p, 100 + N)
    f_gpu = A.inverse(flmn, p)
    f_ref = oracle.inverse(flmn, p, "via_ssht")
    assert _rel(f_gpu, f_ref) < TOL
    rng = np.random.default_rng(7 * N + real)
    f_any = rng.uniform(-1, 1, f_ref.shape)
    if not real:
        f_any = f_any + 1j * rng.uniform(-1, 1, f_ref.shape)
    init = np.zeros(A.flmn_size(p), dtype=np.complex128)
    b_gpu = A.forward(f_any, p, flmn=init.copy())
    b_ref = oracle.forward(f_any, p, "via_ssht", out=init.copy())
    assert _rel(b_gpu, b_ref) < TOL


@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 7, 8, 11, 15, 16, 17, 18, 21, 24, 31, 32, 33, 47, 64])
@pytest.mark.parametrize("real,ss", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_gamma_lengths_vs_oracle(oracle, N, real, ss):
    _check(oracle, N, real, ss)


@pytest.mark.parametrize("N", [25, 33, 38, 41, 53, 64])
@pytest.mark.parametrize("engine", [1, 2])
@pytest.mark.parametrize("real", [0, 1])
def test_gamma_engine_forced_vs_oracle(oracle, tmp_path, N, engine, real):
    t = tmp_path / "t.txt"
    t.write_text("%d %d %d 0 0 0 0 %d\n" % (N + 3, N, real, engine))
    A.tune_load(t)
    try:
        p = P.SO3Parameters(L=N + 3, N=N, reality=real, n_order=P.SO3_N_ORDER_ZERO_FIRST)
        # lengths up to 2N-1 = 63 take the direct DFT whatever the table says
        assert A.tune_get(p)["gamma_engine"] == (0 if 2 * N - 1 <= 63 else engine)
        _check(oracle, N, real, 0)
    finally:
        A.tune_clear()
