"""The MATLAB sample layout of the reference's mex glue
(src/matlab/so3_inverse_mex.c:310-339: f_m[a nG nB + b nG + g] =
f[g nA nB + b nA + a], real and imaginary parts in separate arrays) against
numpy's transpose of the same samples, both directions, bit for bit."""
import numpy as np
import pytest
import torch

import so3_amd.api as A
from so3_amd import params as P

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


@pytest.mark.parametrize("L,N,ss,real,steer", [(16, 16, 0, 0, 0), (33, 5, 1, 0, 0), (20, 7, 0, 1, 0),
                                                (17, 4, 1, 1, 0), (12, 6, 0, 0, 1), (130, 3, 0, 0, 0)])
def test_mex_layout_round_trip(L, N, ss, real, steer):
    p = P.SO3Parameters(L=L, N=N, sampling_scheme=ss, reality=real, steerable=steer)
    na, nb, ng = A.n_alpha(p), A.n_beta(p), A.n_gamma(p)
    rng = np.random.default_rng(L + N)
    f = rng.uniform(-1, 1, na * nb * ng)
    if not real:
        f = f + 1j * rng.uniform(-1, 1, f.size)
    fd = torch.from_numpy(f).to(DEV)
    re = torch.full((f.size,), np.nan, dtype=torch.float64, device=DEV)
    im = torch.full((f.size,), np.nan, dtype=torch.float64, device=DEV)
    A.to_mex_layout_device(re.data_ptr(), 0 if real else im.data_ptr(), fd.data_ptr(), p)
    torch.cuda.synchronize()
    want = f.reshape(ng, nb, na).transpose(2, 1, 0).ravel()
    assert np.array_equal(re.cpu().numpy(), want.real)
    if not real:
        assert np.array_equal(im.cpu().numpy(), want.imag)
    back = torch.zeros_like(fd)
    A.from_mex_layout_device(back.data_ptr(), re.data_ptr(), 0 if real else im.data_ptr(), p)
    torch.cuda.synchronize()
    assert torch.equal(back, fd)
    if not real:  # a missing imaginary array reads as zeros
        A.from_mex_layout_device(back.data_ptr(), re.data_ptr(), 0, p)
        torch.cuda.synchronize()
        assert np.array_equal(back.cpu().numpy(), f.real + 0j)
