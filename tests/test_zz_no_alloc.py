"""GPU: the allocation-free <Z_w> entry point (tqsv_expval_z_all_noalloc)
against the workspace entry point (tqsv_expval_z_all) through the C ABI.

Both run the same two kernels in the same order, so the outputs must be
bit-identical.  Two calls in a row, the second on a smaller input, check that
the fixed scratch array carries nothing over between calls; a shape beyond
the array's maximum must be rejected on the host with a clear error.
"""
import numpy as np
import pytest
import torch

gpu = pytest.mark.gpu
DEV = "cuda"


def _state(n, bsz, seed):
    g = torch.Generator().manual_seed(seed)
    st = torch.randn(bsz, 1 << n, dtype=torch.complex64, generator=g)
    st /= torch.linalg.vector_norm(st, dim=1, keepdim=True)
    return st.to(DEV)


def _with_workspace(st, n, bsz):
    from torchquantum_amd import _native as N
    nbytes = N.lib().tqsv_expval_workspace_bytes(bsz, n, 0)
    ws = torch.empty(max(nbytes // 4, 1), dtype=torch.float32, device=DEV)
    out = torch.empty(bsz, n, dtype=torch.float32, device=DEV)
    N.call("tqsv_expval_z_all", N.stream_of(st), st.data_ptr(), bsz, n, N.fptr(out), ws.data_ptr(), nbytes)
    return out


def _noalloc(st, n, bsz, out):
    from torchquantum_amd import _native as N
    N.call("tqsv_expval_z_all_noalloc", N.stream_of(st), st.data_ptr(), bsz, n, N.fptr(out))
    return out


@gpu
def test_noalloc_matches_workspace_entry_point():
    from oracle import statevector as O
    cases = [(18, 24, 1), (9, 3, 2)]        # second call smaller than the first
    outs = []
    for n, bsz, seed in cases:
        st = _state(n, bsz, seed)
        ref = _with_workspace(st, n, bsz)
        out = torch.full((bsz, n), float("nan"), dtype=torch.float32, device=DEV)
        _noalloc(st, n, bsz, out)
        torch.cuda.synchronize()
        assert torch.equal(out, ref), (n, bsz)
        oracle = O.expval_z_all(st.cpu().numpy().astype(np.complex128), n)
        assert np.abs(out.cpu().numpy() - oracle).max() < 1e-5
        outs.append(out)
    assert outs[0].shape != outs[1].shape


@gpu
def test_noalloc_rejects_oversized_shape():
    from torchquantum_amd import _native as N
    n, bsz = 20, 1 << 17            # needs bsz x (n + 1) floats > 2^20: rejected before any launch
    dummy = torch.empty(8, dtype=torch.complex64, device=DEV)
    out = torch.empty(1, dtype=torch.float32, device=DEV)
    rc = N.lib().tqsv_expval_z_all_noalloc(N.stream_of(dummy), dummy.data_ptr(), bsz, n, N.fptr(out))
    assert rc == -3
    assert b"maximum" in N.lib().tqsv_last_error()
