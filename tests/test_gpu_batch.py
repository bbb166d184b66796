"""GPU: many uneven circuits in level-synchronous batch launches (and their
HIP-graph replay) against the complex128 oracle, 1e-5."""
import numpy as np
import pytest
import torch

from oracle import statevector as O

gpu = pytest.mark.gpu


@gpu
@pytest.mark.parametrize("width", [1, 2, 3])
def test_batch_matches_oracle(width):
    from torchquantum_amd import batch as B
    from torchquantum_amd.engine import EngineOptions
    circs = B.random_circuits(n_circ=10, n_range=(5, 14), depth_range=(2, 8), seed=width)
    cb = B.CircuitBatch(circs, device="cuda", options=[EngineOptions(fusion_width=width)] * len(circs))
    z = cb.run().cpu().numpy()
    for i, (n, ops) in enumerate(circs):
        ref = O.run_ops(ops, n, 1)
        got = cb.state(i).cpu().numpy()
        assert np.abs(got - ref[0]).max() < 1e-5, i
        assert abs(z[i] - O.expval_z_all(ref, n)[0, 0]) < 1e-5


@gpu
def test_batch_graph_replay():
    from torchquantum_amd import batch as B
    circs = B.random_circuits(n_circ=16, n_range=(6, 16), depth_range=(2, 10), seed=9)
    cb = B.CircuitBatch(circs, device="cuda")
    eager = cb.run().clone()
    cb.capture()
    for _ in range(3):
        out = cb.run()
        torch.cuda.synchronize()
        assert torch.allclose(out, eager, atol=1e-6)
        cb.states.zero_()          # eager work between replays
    refs = np.array([O.expval_z_all(O.run_ops(ops, n, 1), n)[0, 0] for n, ops in circs])
    assert np.abs(eager.cpu().numpy() - refs).max() < 1e-5


@gpu
def test_batch_queue_matches_levels():
    """One persistent launch (mode "queue": items taken from a device counter,
    each waiting for its circuit's previous stage) gives bitwise the states
    of one launch per level, with no wait running out, eagerly and in a
    replayed graph."""
    from torchquantum_amd import batch as B
    circs = B.random_circuits(n_circ=24, n_range=(6, 16), depth_range=(2, 12), seed=11)
    cq = B.CircuitBatch(circs, device="cuda", mode="queue")
    cl = B.CircuitBatch(circs, device="cuda", mode="levels")
    zq, zl = cq.run().clone(), cl.run().clone()
    torch.cuda.synchronize()
    cq.check()
    assert torch.equal(cq.states, cl.states)
    assert torch.equal(zq, zl)
    cq.capture()
    for _ in range(3):
        cq.states.zero_()
        out = cq.run()
        torch.cuda.synchronize()
        cq.check()
        assert torch.equal(cq.states, cl.states)
        assert torch.equal(out, zl)


@gpu
def test_batch_queue_wait_that_runs_out_is_reported():
    """The queue's bounded wait, forced to run out from the host: one item is
    given a wait count its circuit never reaches and the poll limit is made
    small, so that wait ends by its limit; the launch still finishes, the
    error flag is set and check() raises. Restored, the next launch is valid
    again (the flag is reset by the launch's first kernel)."""
    from torchquantum_amd import batch as B
    circs = B.random_circuits(n_circ=6, n_range=(6, 12), depth_range=(2, 6), seed=5)
    cb = B.CircuitBatch(circs, device="cuda")
    ref = B.CircuitBatch(circs, device="cuda", mode="levels")
    ref.run()
    q = np.frombuffer(cb.queue.cpu().numpy().tobytes(), dtype=B.QUEUE_ITEM).copy()
    last = len(q) - 1
    saved = int(q["need"][last])
    q["need"][last] = 1 << 30
    cb.queue.copy_(torch.from_numpy(q.view(np.uint8).copy()))
    cb.spin_limit = 1000
    cb.run()
    torch.cuda.synchronize()
    with pytest.raises(RuntimeError, match="ran out"):
        cb.check()
    q["need"][last] = saved
    cb.queue.copy_(torch.from_numpy(q.view(np.uint8).copy()))
    cb.run()
    torch.cuda.synchronize()
    cb.check()
    assert torch.equal(cb.states, ref.states)
