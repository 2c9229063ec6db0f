"""Serving loop: results of requests served in a mix (continuous batching,
concurrent lanes and caller streams) are bit-identical to the same requests
run alone through the plain entry points."""
import numpy as np
import pytest

import so3_amd.api as A
from so3_amd import params as P
from so3_amd import serve

pytestmark = pytest.mark.gpu


def _mix():
    rng = np.random.default_rng(9)
    reqs = []
    t = 0.0
    for i in range(14):
        t += rng.exponential(0.004)
        if i % 3 == 2:
            L, N, kind = int(rng.integers(16, 48)), int(rng.choice([4, 8])), A.ITEM_S2TOSO3_INVERSE
        else:
            L = int(rng.integers(8, 40))
            N, kind = int(rng.integers(1, min(L, 12) + 1)), A.ITEM_PAIR
        reqs.append(serve.Request(i, kind, P.SO3Parameters(L=L, N=N, n_order=P.SO3_N_ORDER_ZERO_FIRST), t))
    return reqs


def _alone(r, fn):
    d = fn(r)
    if r.kind == A.ITEM_PAIR:
        f = A.inverse(d["flmn"], r.p)
        return {"f": f, "flmn_out": A.forward(f, r.p)}
    h = A.s2toso3(d["flm"], d["glm"], r.p)
    return {"f": A.inverse(h, r.p), "flmn_out": h}


@pytest.mark.parametrize("mode", [A.BATCH_PERSISTENT, A.BATCH_LANES])
@pytest.mark.parametrize("budget,max_batch", [(8 << 30, 8), (1, 3)])
def test_served_equals_alone(budget, max_batch, mode):
    reqs = _mix()
    fn = serve.random_inputs(1)
    srv = serve.Server(hbm_budget=budget, lanes=3, max_batch=max_batch, streams=2, mode=mode)
    srv.prepare(reqs)
    srv.run(reqs, fn, keep_outputs=True)
    for r in reqs:
        assert r.done is not None and r.done >= r.arrival
        ref = _alone(r, fn)
        for k, v in ref.items():
            assert np.array_equal(r.outputs[k], v), (r.rid, k)
    st = serve.latency_stats(reqs)
    assert st["requests"] == len(reqs) and st["p99_ms"] >= st["p50_ms"] >= 0
