"""Delta tables are shared across L: a table built for L' serves every
L <= L'.  A transform run on a larger table is bit-identical to the same
transform on its own table (both recursions, complex and real, MW and
MW_SS, inverse and forward), and a larger L builds one new table."""
import numpy as np
import pytest

import so3_amd.api as A
from so3_amd import params as P
from so3_amd import synth
from so3_amd._lib import lib

pytestmark = pytest.mark.gpu


def _run(p, flmn):
    f = A.inverse(flmn, p)
    return f, A.forward(f, p)


@pytest.mark.parametrize("dl", [0, 1])
@pytest.mark.parametrize("L,N,real,ss", [(40, 17, 0, 0), (37, 37, 1, 1), (129, 9, 0, 1), (250, 3, 1, 0)])
def test_shared_table_is_bit_identical(dl, L, N, real, ss):
    p = P.SO3Parameters(L=L, N=N, reality=real, sampling_scheme=ss, dl_method=dl,
                        n_order=P.SO3_N_ORDER_ZERO_FIRST)
    flmn = synth.gen_flmn(p, 5)
    assert lib().so3x_release_all() == 0
    own = _run(p, flmn)                       # builds the L table
    assert lib().so3x_release_all() == 0
    A.prepare(p.replace(L=L + 77, N=2))       # only a larger table exists now
    shared = _run(p, flmn)
    for a, b in zip(own, shared):
        assert np.array_equal(a, b)
    assert lib().so3x_release_all() == 0
