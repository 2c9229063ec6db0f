"""The persistent batch mode (so3x_batch_persistent, SO3X_BATCH_PERSISTENT
graphs): the contractions of all items run in one launch per direction and
signal type, fed by device queues, the DFT stages per item on lanes.  Every
item must produce exactly the bits of the same transform run alone, eagerly
and over graph replays, at any lane count; the tuner's mode choice is
reported and used by AUTO batches."""
import pytest
import torch

import so3_amd.api as A
from tests.test_gpu_batch import _check, _clear, _setup

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("lanes", [1, 3, 8])
def test_persistent_batch_matches_alone(lanes):
    items, refs, outs = _setup()
    _clear(outs)
    A.batch_persistent(items, lanes=lanes, stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    _check(refs, outs)


def test_persistent_batches_back_to_back():
    """two eager persistent batches enqueued without a host wait share the
    device's buffers in stream order"""
    items, refs, outs = _setup()
    s = torch.cuda.current_stream().cuda_stream
    _clear(outs)
    A.batch_persistent(items, lanes=2, stream=s)
    A.batch_persistent(items, lanes=4, stream=s)
    torch.cuda.synchronize()
    _check(refs, outs)


def test_persistent_graph_replays():
    items, refs, outs = _setup()
    g = A.graph_create(items, lanes=4, mode=A.BATCH_PERSISTENT)
    try:
        s = torch.cuda.current_stream()
        for rep in range(3):
            _clear(outs)
            torch.cuda.synchronize()
            A.graph_launch(g, s.cuda_stream)
            x = torch.ones(1 << 20, dtype=torch.float64, device="cuda:0") * rep
            torch.cuda.synchronize()
            assert float(x.sum()) == rep * (1 << 20)
            _check(refs, outs)
    finally:
        A.graph_destroy(g)


def test_tuner_mode_choice():
    items, refs, outs = _setup()
    s = torch.cuda.current_stream().cuda_stream
    lanes = A.batch_tune(items, stream=s)
    mode, l2 = A.batch_choice(items)
    assert mode in (A.BATCH_LANES, A.BATCH_PERSISTENT) and l2 == lanes
    _clear(outs)
    A.batch(items, lanes=0, stream=s)  # AUTO: the tuned mode
    torch.cuda.synchronize()
    _check(refs, outs)
