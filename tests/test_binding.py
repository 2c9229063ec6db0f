"""The Python binding's host helpers (so3_amd.api, mirroring
src/so3/bindings.pyx:96-448) without a GPU: parameter construction from enum
names, the n / l / m loop generators against the reference's loop-range
tests (tests/test_convolution.c:34-132), is_elmn_non_zero, and the
convolved parameters (tests/test_convolution.c:191-213)."""
import pytest

import so3_amd.api as so3
import so3_amd.params as P


def test_create_parameter_dict_defaults_and_names():
    p = so3.create_parameter_dict(16, 8)
    assert (p.L, p.N, p.L0, p.reality, p.steerable) == (16, 8, 0, 0, 0)
    assert p.sampling_scheme == P.SO3_SAMPLING_MW
    assert p.n_order == P.SO3_N_ORDER_NEGATIVE_FIRST       # the binding's default
    assert p.storage == P.SO3_STORAGE_PADDED
    assert p.n_mode == P.N_MODE_ALL and p.dl_method == P.SSHT_DL_RISBO
    q = so3.create_parameter_dict(9, 3, 2, reality=1, sampling_scheme_str="SO3_SAMPLING_MW_SS",
                                  n_order_str="SO3_N_ORDER_ZERO_FIRST", storage_str="SO3_STORAGE_COMPACT",
                                  n_mode_str="SO3_N_MODE_ODD", dl_method_str="SSHT_DL_TRAPANI", steerable=0)
    assert (q.sampling_scheme, q.n_order, q.storage, q.n_mode, q.dl_method) == (1, 0, 1, 2, 1)
    # a name the binding does not know selects its else branch
    r = so3.create_parameter_dict(4, 2, n_mode_str="bogus", storage_str="bogus")
    assert r.n_mode == P.N_MODE_L and r.storage == P.SO3_STORAGE_COMPACT
    d = {k: getattr(q, k) for k, _ in P.so3_parameters_t._fields_}
    assert vars(P.SO3Parameters().from_dict(d)) == vars(q)


@pytest.mark.parametrize("N", [1, 2, 5, 8])
@pytest.mark.parametrize("real", [0, 1])
def test_n_loop_ranges(N, real):
    for mode in range(5):
        p = P.SO3Parameters(L=10, N=N, reality=real, n_mode=mode)
        top_even = (N - 1) % 2 == 0
        if real:
            want = {0: (0, N - 1, 1), 4: (0, N - 1, 1),
                    1: (0, N - 1 if top_even else N - 2, 2),
                    2: (1, N - 1 if not top_even else N - 2, 2),
                    3: (N - 1, N - 1, 1)}[mode]
        else:
            want = {0: (-N + 1, N - 1, 1), 4: (-N + 1, N - 1, 1),
                    1: ((-N + 1) if top_even else (-N + 2), (N - 1) if top_even else (N - 2), 2),
                    2: ((-N + 1) if not top_even else (-N + 2), (N - 1) if not top_even else (N - 2), 2),
                    3: (-N + 1, N - 1, 2 * N - 2 if N > 1 else 1)}[mode]
        assert so3.get_loop_n_values(p) == want
        assert list(so3.loop_over_n(p)) == list(range(want[0], want[1] + 1, want[2]))


def test_el_and_m_loops():
    p = P.SO3Parameters(L=12, N=5, L0=3)
    assert so3.get_loop_el_values(-2, p) == (3, 11, 1)
    assert so3.get_loop_el_values(4, p) == (4, 11, 1)
    assert list(so3.loop_over_el(4, p)) == list(range(4, 12))
    q = p.replace(n_mode=P.N_MODE_L, L0=0)
    assert so3.get_loop_el_values(-3, q) == (3, 3, 1)       # N_MODE_L keeps l = |n|
    assert list(so3.loop_over_el(0, q)) == [0]
    assert so3.get_loop_m_values(4) == (-4, 4, 1)
    assert list(so3.loop_over_m(2)) == [-2, -1, 0, 1, 2]


def test_is_elmn_non_zero():
    p = P.SO3Parameters(L=8, N=4, L0=2, n_mode=P.N_MODE_EVEN)
    assert so3.is_elmn_non_zero(3, 1, 2, p)
    assert not so3.is_elmn_non_zero(3, 1, 1, p)      # odd n outside the even n-mode
    assert not so3.is_elmn_non_zero(1, 0, 0, p)      # l < L0
    assert not so3.is_elmn_non_zero(3, 4, 2, p)      # |m| > l
    assert not so3.is_elmn_non_zero(1, 0, 2, P.SO3Parameters(L=8, N=4))  # l < |n|


def test_get_convolved_parameters_known_answer():
    f = P.SO3Parameters(L0=1, L=3, N=2, n_order=0)
    g = P.SO3Parameters(L0=0, L=4, N=4, n_order=0)
    h = so3.get_convolved_parameters(f, g)
    assert (h.L, h.L0, h.N) == (3, 1, 3)
    hd = so3.get_convolved_parameter_dict({k: getattr(f, k) for k, _ in P.so3_parameters_t._fields_},
                                          {k: getattr(g, k) for k, _ in P.so3_parameters_t._fields_})
    assert vars(hd) == vars(h)
