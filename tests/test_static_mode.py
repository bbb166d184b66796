"""Static mode (QuantumModule.static_on / static_support, modules.py:230-316,
graphs.py:60-81) routed to the fused planner.

Host tests: flag propagation, block-width options, the flush at the end of
the graph-top forward.  GPU tests follow the reference's
test/static_mode_test.py:258-401 method: the same module run dynamically and in
static mode with every wires_per_block; states agree to 1e-5 and gradients to
1e-4 (the reference's tolerances)."""
import numpy as np
import pytest
import torch

import torchquantum_amd as tq
import torchquantum_amd.functional as tqf

gpu = pytest.mark.gpu


class QLayer(tq.QuantumModule):
    def __init__(self, n_wires, seed):
        super().__init__()
        self.n_wires = n_wires
        self.rd = tq.RandomLayer(wires=list(range(n_wires)), n_ops=40, seed=seed,
                                 op_types=(tq.RX, tq.RY, tq.RZ, tq.U3, tq.CNOT, tq.CZ, tq.SWAP, tq.CRX, tq.RZZ,
                                           tq.Hadamard, tq.Toffoli))
        self.rx = tq.RX(has_params=True, trainable=True)
        self.cu3 = tq.CU3(has_params=True, trainable=True)

    @tq.static_support
    def forward(self, q_device):
        self.rd(q_device)
        for w in range(self.n_wires):
            self.rx(q_device, wires=w)
        for w in range(self.n_wires - 1):
            self.cu3(q_device, wires=[w, w + 1])
        tqf.h(q_device, wires=0, static=self.static_mode, parent_graph=self.graph)


def test_static_on_propagates():
    m = QLayer(4, seed=0)
    m.static_on(wires_per_block=3)
    assert m.static_mode and m.is_graph_top and m.wires_per_block == 3
    assert m.rd.static_mode and not m.rd.is_graph_top
    assert all(o.static_mode for o in m.rd.op_list)
    assert m.static_options().mode == "fused" and m.static_options().fusion_width == 3
    m.static_on(wires_per_block=9)
    assert m.static_options().fusion_width == 6        # widest block the planner builds
    m.static_off()
    assert not m.static_mode and not m.rd.static_mode
    with pytest.raises(ValueError):
        m.static_on(wires_per_block=0)


def test_static_forward_flushes_with_block_width(monkeypatch):
    """The graph-top forward ends with one flush carrying the static options;
    sub-modules (not graph-top) do not flush; dynamic mode does not flush."""
    m = QLayer(3, seed=1)
    q = tq.QuantumDevice(n_wires=3, bsz=1, device="cpu")
    seen = []
    monkeypatch.setattr(q, "flush", lambda options=None: seen.append(options))
    m(q)
    assert seen == []
    m.static_on(wires_per_block=2)
    m(q)
    assert len(seen) == 1 and seen[0].fusion_width == 2 and seen[0].mode == "fused"
    assert len(q._tape) > 0      # the gates were recorded on the tape (the flush was stubbed out)


@gpu
@pytest.mark.parametrize("n_wires", [2, 3, 5, 8])
def test_static_vs_dynamic_states(n_wires):
    torch.manual_seed(42)
    m = QLayer(n_wires, seed=n_wires).cuda()
    bsz = 3
    q0 = tq.QuantumDevice(n_wires=n_wires, bsz=bsz, device="cuda")
    q0.engine_options = tq.EngineOptions(mode="pergate")
    m(q0)
    ref = q0.get_states_1d().detach().clone()
    for wpb in range(1, n_wires + 1):
        q1 = tq.QuantumDevice(n_wires=n_wires, bsz=bsz, device="cuda")
        m.static_off()
        m.static_on(wires_per_block=wpb)
        m(q1)
        assert not q1._tape      # flushed at the end of the graph-top forward
        got = q1.get_states_1d().detach()
        assert torch.allclose(ref, got, atol=1e-5), (n_wires, wpb, float((ref - got).abs().max()))
        m.static_off()


@gpu
@pytest.mark.parametrize("n_wires", [3, 6])
def test_static_vs_dynamic_gradients(n_wires):
    torch.manual_seed(7)
    m = QLayer(n_wires, seed=10 + n_wires).cuda()
    bsz = 2

    def grads(static_wpb):
        q = tq.QuantumDevice(n_wires=n_wires, bsz=bsz, device="cuda")
        m.static_off()
        if static_wpb:
            m.static_on(wires_per_block=static_wpb)
        else:
            q.engine_options = tq.EngineOptions(mode="pergate")
        m(q)
        loss = q.get_states_1d().abs().sum()
        loss.backward()
        g = {k: p.grad.detach().clone() for k, p in m.named_parameters() if p.grad is not None}
        m.zero_grad(set_to_none=True)
        m.static_off()
        return g

    g0 = grads(0)
    assert len(g0) > 0
    for wpb in range(1, n_wires + 1):
        g1 = grads(wpb)
        assert g1.keys() == g0.keys()
        for k in g0:
            assert torch.allclose(g0[k], g1[k], atol=1e-4), (n_wires, wpb, k)
