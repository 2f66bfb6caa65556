"""szsnn_adj_input_grad_noalloc (the input gradient with its scratch in the
library's own device array, no allocation) against szsnn_adj_input_grad
(workspace argument): bit-identical results on two calls in a row, the
second with a smaller input, and a clear error beyond the array's size."""
import ctypes

import pytest

pytestmark = pytest.mark.gpu


def _case(torch, cuda, rows, n_in, H, seed):
    g = torch.Generator().manual_seed(seed)
    X = (torch.rand(rows, n_in, generator=g) < 0.03).float() * (1.0 + (torch.rand(rows, n_in, generator=g) < 0.2))
    W = torch.randn(n_in, H, generator=g) * 0.1
    LE = torch.randn(rows, H, generator=g)
    LI = torch.randn(rows, H, generator=g)
    return [t.to(cuda).contiguous() for t in (X, W, LE, LI)]


def test_noalloc_input_grad_equals_workspace_entry():
    import torch
    from szsnn_amd import _lib
    cuda = torch.device("cuda:0")
    lib = _lib.load()
    ptr = lambda t: ctypes.c_void_p(t.data_ptr())
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    for rows, n_in, H, seed in ((12800, 1000, 2000, 1), (300, 77, 2100, 2)):   # the second call is smaller
        print(f"case: no-alloc input grad rows={rows} n_in={n_in} H={H}", flush=True)
        X, W, LE, LI = _case(torch, cuda, rows, n_in, H, seed)
        nbytes = lib.szsnn_adj_input_grad_workspace(rows, n_in)
        ws = torch.empty(nbytes // 8, dtype=torch.int64, device=cuda)
        ref = torch.full((n_in, H), float("nan"), device=cuda)
        out = torch.full((n_in, H), float("nan"), device=cuda)
        assert lib.szsnn_adj_input_grad(stream, rows, n_in, H, ptr(X), ptr(W), ptr(LE), ptr(LI), ptr(ref), ptr(ws),
                                        nbytes) == 0
        assert lib.szsnn_adj_input_grad_noalloc(stream, rows, n_in, H, ptr(X), ptr(W), ptr(LE), ptr(LI),
                                                ptr(out)) == 0
        torch.cuda.synchronize()
        assert not torch.isnan(out).any()
        assert torch.equal(out, ref)
    # beyond the scratch array (2^21 words, 16 MiB): rejected before any launch
    X, W, LE, LI = _case(torch, cuda, 2, 3, 4, 3)
    assert lib.szsnn_adj_input_grad_workspace(64 * 2049, 1024) > 16 << 20
    rc = lib.szsnn_adj_input_grad_noalloc(stream, 64 * 2049, 1024, 4, ptr(X), ptr(W), ptr(LE), ptr(LI), ptr(W))
    assert rc == _lib.SZSNN_ERR_INVALID
    assert b"16 MiB" in lib.szsnn_last_error(None)
