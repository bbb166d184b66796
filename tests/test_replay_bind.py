"""CPU: how a circuit replay follows its inputs (torchquantum_amd.replay.
bind_inputs).  Two recordings of the same function with different inputs:
views of an input need nothing, copies of input columns (GeneralEncoder's
advanced indexing) are found column by column, parameters that do not depend
on the inputs are left alone, and anything else is rejected."""
import pytest
import torch

import torchquantum_amd as tq
import torchquantum_amd.functional as tqf
from torchquantum_amd.replay import bind_inputs

N, BSZ = 5, 3


def _tape(record, *inputs):
    q = tq.QuantumDevice(n_wires=N, bsz=BSZ, device="cpu")
    record(q, *inputs)
    return list(q._tape)


def _two(record, shapes, seed=0):
    g = torch.Generator().manual_seed(seed)
    xa = [torch.rand(*s, generator=g) for s in shapes]
    xb = [torch.rand(*s, generator=g) for s in shapes]
    return bind_inputs(_tape(record, *xa), _tape(record, *xb), xa, xb, BSZ)


def test_views_copies_and_constants():
    enc = tq.GeneralEncoder([{"input_idx": [3], "func": "ry", "wires": [0]},
                             {"input_idx": [0], "func": "rx", "wires": [1]},
                             {"input_idx": [1, 4, 2], "func": "u3", "wires": [2]}])
    theta = torch.tensor([[0.3]], requires_grad=True)
    fixed = torch.tensor([[0.7]])

    def record(q, x, y):
        enc(q, x)                                   # copies of columns 3, 0 and (1, 4, 2) of x
        tqf.rz(q, 3, params=y[:, 1:2])              # a view of y
        tqf.ry(q, 4, params=x[:, 2])                # a view of x
        tqf.rx(q, 0, params=theta)                  # trainable, input-independent
        tqf.cnot(q, [0, 1])
        tqf.phaseshift(q, 2, params=fixed)          # constant
        tqf.ry(q, 1, params=y[:, [0]])              # a copy of a column of the second input

    binds = _two(record, [(BSZ, 5), (BSZ, 2)])
    assert binds == [(0, 0, [3]), (1, 0, [0]), (2, 0, [1, 4, 2]), (8, 1, [0])]


def test_transformed_input_rejected():
    def record(q, x):
        tqf.ry(q, 0, params=x[:, [0]] * 2.0)        # an arithmetic transform of a column
    with pytest.raises(ValueError, match="neither a view nor a copy"):
        _two(record, [(BSZ, 2)])


def test_structure_change_rejected():
    def record(q, x):
        tqf.ry(q, 0, params=x[:, 0:1])
        if float(x[0, 1]) > 0.5:
            tqf.cnot(q, [0, 1])
        else:
            tqf.cz(q, [0, 1])
    with pytest.raises(ValueError, match="fixed structure"):
        g = torch.Generator().manual_seed(0)
        xa = torch.rand(BSZ, 2, generator=g)
        xb = xa.clone()
        xa[0, 1], xb[0, 1] = 0.9, 0.1
        bind_inputs(_tape(record, xa), _tape(record, xb), [xa], [xb], BSZ)


def test_many_encoder_columns_bind_quickly():
    """An encoder with hundreds of input columns (an image-sized input):
    each gate's column is found by a host-side index, not by comparing
    against every column."""
    import time
    n_feat = 400
    enc = tq.GeneralEncoder([{"input_idx": [(7 * i) % n_feat], "func": "ry", "wires": [i % N]}
                             for i in range(n_feat)])
    t0 = time.perf_counter()
    binds = _two(lambda q, x: enc(q, x), [(BSZ, n_feat)], seed=2)
    assert time.perf_counter() - t0 < 20.0
    assert binds == [(i, 0, [(7 * i) % n_feat]) for i in range(n_feat)]
