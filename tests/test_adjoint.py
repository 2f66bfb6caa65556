"""Discrete adjoint of the feedforward IF_curr_exp network.

CPU: the FP64 torch restatement (oracle/adjoint_ref.py) is pinned to the C++
oracle's IF_curr_exp trace, and its fixed-spike autograd gradient to central
finite differences of the free-running forward (perturbations checked not to
flip any spike).
GPU: the HIP adjoint's dL/dW1, dL/dW2 against the FP64 autograd reference
with the GPU's spike pattern held fixed (relative L2 error <= 1e-4), and
against central finite differences.
"""
import numpy as np
import pytest
import torch

REL_TOL = 1e-4


def _to64(P, keys=("W1", "W2", "X", "target", "params")):
    return {k: torch.tensor(P[k], dtype=torch.float64) for k in keys}


def test_fp64_reference_matches_oracle_neuron_trace(oracle_mod):
    from oracle import adjoint_ref
    from szsnn_amd.adjoint import make_problem
    P = make_problem(batch=3, n_in=50, n_hidden=20, n_out=4, steps=80, w1_mean=0.3, w1_std=0.5)
    d = _to64(P)
    _, sp, v = adjoint_ref.forward(d["W1"], d["W2"], d["X"], d["target"], d["params"], P["alpha"],
                                   return_v=True)
    assert sp.any()
    W1p = d["W1"].clamp(min=0)
    W1n = d["W1"].clamp(max=0)
    for b, j in ((0, 0), (1, 7), (2, 19)):
        xprev = torch.cat([torch.zeros(1, 50, dtype=torch.float64), d["X"][:-1, b]], 0)
        ae = (xprev @ W1p[:, j]).numpy()
        ai = (xprev @ W1n[:, j]).numpy()
        tr, spk = oracle_mod.neuron_trace(0, 64, P["params"][:, j].astype(np.float64),
                                          [P["params"][0, j], 0, 0, 10000.0], ae, ai)
        np.testing.assert_allclose(v[:, b, j].numpy(), tr[:, 0], rtol=0, atol=1e-10)
        np.testing.assert_array_equal(sp[:, b, j].numpy(), spk)


def test_fp64_autograd_matches_finite_differences():
    from oracle import adjoint_ref
    from szsnn_amd.adjoint import make_problem
    P = make_problem(batch=4, n_in=60, n_hidden=30, n_out=5, steps=60, w1_mean=0.3, w1_std=0.5)
    d = _to64(P)
    W1 = d["W1"].clone().requires_grad_(True)
    W2 = d["W2"].clone().requires_grad_(True)
    loss, sp = adjoint_ref.forward(W1, W2, d["X"], d["target"], d["params"], P["alpha"])
    assert sp.any() and not sp.all()
    loss_fixed, _ = adjoint_ref.forward(W1, W2, d["X"], d["target"], d["params"], P["alpha"], spikes=sp)
    loss_fixed.backward()
    rng = np.random.default_rng(0)
    checked = 0
    for _ in range(40):
        which = rng.integers(0, 2)
        W = d["W1"] if which == 0 else d["W2"]
        i, j = (int(rng.integers(0, W.shape[0])), int(rng.integers(0, W.shape[1])))
        eps = 1e-6
        vals = []
        for sgn in (1, -1):
            Wp = W.clone()
            Wp[i, j] += sgn * eps
            args = (Wp, d["W2"]) if which == 0 else (d["W1"], Wp)
            l, s2 = adjoint_ref.forward(*args, d["X"], d["target"], d["params"], P["alpha"])
            if not torch.equal(s2, sp):
                break
            vals.append(float(l))
        if len(vals) < 2:
            continue  # a spike flipped: not a valid finite-difference point
        fd = (vals[0] - vals[1]) / (2 * eps)
        g = float((W1.grad if which == 0 else W2.grad)[i, j])
        assert abs(fd - g) <= 1e-6 * max(1.0, abs(g)) + 1e-7, (which, i, j, fd, g)
        checked += 1
    assert checked >= 20


@pytest.fixture(scope="module")
def cuda():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


def _gpu_grads(P, device, K=20):
    from szsnn_amd.adjoint import snn_loss
    t = {k: torch.tensor(P[k], dtype=torch.float32, device=device).contiguous()
         for k in ("W1", "W2", "X", "target", "params")}
    W1 = t["W1"].clone().requires_grad_(True)
    W2 = t["W2"].clone().requires_grad_(True)
    loss, y, bits = snn_loss(W1, W2, t["X"], t["target"], t["params"], P["alpha"], K=K)
    loss.backward()
    torch.cuda.synchronize()
    return float(loss.detach()), W1.grad.cpu().double(), W2.grad.cpu().double(), bits.cpu()


def _bits_to_spikes(bits, T, B, H):
    b = bits.numpy().view(np.uint64)
    u = np.unpackbits(b.view(np.uint8).reshape(T, -1), axis=1, bitorder="little")[:, :B * H]
    return torch.tensor(u.reshape(T, B, H).astype(bool))


def _ref_grads(P, spikes):
    from oracle import adjoint_ref
    d = _to64(P)
    W1 = d["W1"].clone().requires_grad_(True)
    W2 = d["W2"].clone().requires_grad_(True)
    loss, _ = adjoint_ref.forward(W1, W2, d["X"], d["target"], d["params"], P["alpha"], spikes=spikes)
    loss.backward()
    return float(loss), W1.grad, W2.grad


def _rel(a, b):
    return float((a - b).norm() / max(float(b.norm()), 1e-30))


@pytest.mark.gpu
@pytest.mark.parametrize("variant", ["spec", "spiking"])
def test_hip_adjoint_matches_fp64_autograd(cuda, variant):
    from szsnn_amd.adjoint import make_problem
    if variant == "spec":
        P = make_problem()  # B=64, 1000 -> 2000 -> 100, T=200, W1 ~ N(0, 0.02)
    else:
        P = make_problem(batch=16, n_hidden=1000, w1_mean=0.02, w1_std=0.1)
    T, B, _ = P["X"].shape
    H = P["W1"].shape[1]
    print(f"case: adjoint {variant} T={T} B={B} H={H}", flush=True)
    loss, g1, g2, bits = _gpu_grads(P, cuda)
    spikes = _bits_to_spikes(bits, T, B, H)
    if variant == "spiking":
        assert spikes.float().mean() > 0.02
    rloss, r1, r2 = _ref_grads(P, spikes)
    assert abs(loss - rloss) <= 1e-5 * abs(rloss)
    assert _rel(g1, r1) <= REL_TOL, _rel(g1, r1)
    assert _rel(g2, r2) <= REL_TOL, _rel(g2, r2)


@pytest.mark.gpu
def test_hip_adjoint_checkpoint_interval_invariant(cuda):
    from szsnn_amd.adjoint import make_problem
    P = make_problem(batch=8, n_hidden=500, steps=90, w1_mean=0.02, w1_std=0.1)
    _, a1, a2, _ = _gpu_grads(P, cuda, K=20)
    _, b1, b2, _ = _gpu_grads(P, cuda, K=7)
    _, c1, c2, _ = _gpu_grads(P, cuda, K=64)
    # the recurrences are identical; only the grouping of the per-segment GEMM
    # sums into dW changes with K, at FP32 rounding level
    for x, y in ((a1, b1), (a1, c1), (a2, b2), (a2, c2)):
        assert _rel(x, y) <= 1e-6, _rel(x, y)


@pytest.mark.gpu
def test_hip_adjoint_matches_finite_differences(cuda):
    from oracle import adjoint_ref
    from szsnn_amd.adjoint import make_problem
    P = make_problem(batch=8, n_hidden=300, n_in=200, n_out=20, steps=100, w1_mean=0.05, w1_std=0.2)
    _, g1, g2, bits = _gpu_grads(P, cuda)
    d = _to64(P)
    _, sp = adjoint_ref.forward(d["W1"], d["W2"], d["X"], d["target"], d["params"], P["alpha"])
    rng = np.random.default_rng(3)
    checked = 0
    for _ in range(30):
        which = int(rng.integers(0, 2))
        W = d["W1"] if which == 0 else d["W2"]
        i, j = int(rng.integers(0, W.shape[0])), int(rng.integers(0, W.shape[1]))
        eps = 1e-5
        vals = []
        for sgn in (1, -1):
            Wp = W.clone()
            Wp[i, j] += sgn * eps
            args = (Wp, d["W2"]) if which == 0 else (d["W1"], Wp)
            l, s2 = adjoint_ref.forward(*args, d["X"], d["target"], d["params"], P["alpha"])
            if not torch.equal(s2, sp):
                break
            vals.append(float(l))
        if len(vals) < 2:
            continue
        fd = (vals[0] - vals[1]) / (2 * eps)
        g = float((g1 if which == 0 else g2)[i, j])
        assert abs(fd - g) <= 1e-3 * max(abs(fd), 1e-3), (which, i, j, fd, g)
        checked += 1
    assert checked >= 15


@pytest.mark.gpu
@pytest.mark.parametrize("rows,n_in,H", [(300, 1000, 2000), (37, 300, 2100), (5, 513, 4500)])
def test_input_currents_equal_sign_split_products(cuda, rows, n_in, H):
    """szsnn_adj_input_currents (nonzero inputs only, ascending order) against
    the FP64 products X max(W1, 0) and X min(W1, 0); inputs include values
    other than 0/1, widths not multiples of the block (H up to 4500: several
    column passes) and an all-zero row."""
    import ctypes
    import torch
    from szsnn_amd import _lib
    lib = _lib.load()
    g = torch.Generator().manual_seed(rows * 7 + H)
    # values of 2 in the inputs of every other 64-input tile (the kernel reads
    # X in tiles that hold a value other than 0 or 1 and takes 1 elsewhere)
    two = (torch.rand(rows, n_in, generator=g) < 0.2) & ((torch.arange(n_in) % 128) < 64)
    X = (torch.rand(rows, n_in, generator=g) < 0.03).float() * (1.0 + two)
    X[rows // 2] = 0.0
    W = torch.randn(n_in, H, generator=g) * 0.1
    Xd, Wd = X.to(cuda).contiguous(), W.to(cuda).contiguous()
    AE = torch.empty(rows, H, device=cuda)
    AI = torch.empty(rows, H, device=cuda)
    ptr = lambda t: ctypes.c_void_p(t.data_ptr())
    rc = lib.szsnn_adj_input_currents(ctypes.c_void_p(torch.cuda.current_stream().cuda_stream), rows, n_in, H,
                                      ptr(Xd), ptr(Wd), ptr(AE), ptr(AI))
    assert rc == 0
    torch.cuda.synchronize()
    X64, W64 = X.double(), W.double()
    re, ri = X64 @ W64.clamp(min=0), X64 @ W64.clamp(max=0)
    # FP32 sums of at most ~3 % of n_in terms: within a few ulps of the FP64 value
    assert torch.allclose(AE.cpu().double(), re, rtol=1e-5, atol=1e-6)
    assert torch.allclose(AI.cpu().double(), ri, rtol=1e-5, atol=1e-6)
    assert float(AE[rows // 2].abs().max()) == 0.0 and float(AI[rows // 2].abs().max()) == 0.0


@pytest.mark.gpu
@pytest.mark.parametrize("rows,n_in,H", [(12800, 1000, 2000), (300, 77, 2100), (65, 513, 4500), (1, 5, 3)])
def test_input_grad_equals_sign_selected_products(cuda, rows, n_in, H):
    """szsnn_adj_input_grad (rows where the input spiked only, ascending)
    against the FP64 products X^T LE and X^T LI selected by the weight's
    sign, and equal to the same FP32 operations in ascending row order;
    inputs with values other than 0/1, an input that never spikes, row
    counts not multiples of 64, widths beyond one column chunk, zero weights
    (sign >= 0: the E adjoint)."""
    import ctypes
    import torch
    from szsnn_amd import _lib
    lib = _lib.load()
    g = torch.Generator().manual_seed(rows * 3 + H)
    # values of 2 in the inputs of every other 64-input tile (the kernel reads
    # X in tiles that hold a value other than 0 or 1 and takes 1 elsewhere)
    two = (torch.rand(rows, n_in, generator=g) < 0.2) & ((torch.arange(n_in) % 128) < 64)
    X = (torch.rand(rows, n_in, generator=g) < 0.03).float() * (1.0 + two)
    X[:, n_in // 2] = 0.0
    W = torch.randn(n_in, H, generator=g) * 0.1
    W[0, : min(H, 7)] = 0.0
    LE = torch.randn(rows, H, generator=g)
    LI = torch.randn(rows, H, generator=g)
    d = {k: v.to(cuda).contiguous() for k, v in (("X", X), ("W", W), ("LE", LE), ("LI", LI))}
    nbytes = lib.szsnn_adj_input_grad_workspace(rows, n_in)
    assert nbytes == (n_in + (n_in + 63) // 64) * ((rows + 63) // 64) * 8  # bits, and a word per 64 x 64 tile
    ws = torch.empty(nbytes // 8, dtype=torch.int64, device=cuda)
    dW = torch.full((n_in, H), float("nan"), device=cuda)
    ptr = lambda t: ctypes.c_void_p(t.data_ptr())
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert lib.szsnn_adj_input_grad(stream, rows, n_in, H, ptr(d["X"]), ptr(d["W"]), ptr(d["LE"]), ptr(d["LI"]),
                                    ptr(dW), ptr(ws), nbytes - 1) == -1   # workspace too small
    rc = lib.szsnn_adj_input_grad(stream, rows, n_in, H, ptr(d["X"]), ptr(d["W"]), ptr(d["LE"]), ptr(d["LI"]),
                                  ptr(dW), ptr(ws), nbytes)
    assert rc == 0
    torch.cuda.synchronize()
    X64 = X.double()
    ref = torch.where(W >= 0, X64.t() @ LE.double(), X64.t() @ LI.double())
    # FP32 sums over the ~3 % spiking rows: relative to the scale of the sum
    scale = (X64.t() @ LE.double().abs()).clamp(min=1.0)
    err = ((dW.cpu().double() - ref).abs() / scale).max()
    assert float(err) <= 1e-5, float(err)
    assert float(dW[n_in // 2].abs().max()) == 0.0
    # the same FP32 operations in the same order: per (input, column), the
    # products x * adjoint of the spiking rows added in ascending row order
    import numpy as np
    Xn, Wn, LEn, LIn = X.numpy(), W.numpy(), LE.numpy(), LI.numpy()
    got = dW.cpu().numpy()
    for i in range(n_in):
        rr = np.flatnonzero(Xn[:, i])
        if rr.size == 0:
            continue
        terms = Xn[rr, i][:, None] * np.where(Wn[i] >= 0, LEn[rr], LIn[rr])
        want = np.add.accumulate(terms, axis=0, dtype=np.float32)[-1]
        np.testing.assert_array_equal(got[i], want, err_msg=f"input {i}")


@pytest.mark.gpu
def test_input_grad_rejects_rows_beyond_the_bit_kernel_grid(cuda):
    """More than 65535 * 64 rows: both entry points return SZSNN_ERR_INVALID
    with a message, before any launch."""
    import ctypes
    import torch
    from szsnn_amd import _lib
    lib = _lib.load()
    rows, n_in, H = 65535 * 64 + 1, 3, 2
    t = torch.zeros(16, device=cuda)
    p = ctypes.c_void_p(t.data_ptr())
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert lib.szsnn_adj_input_grad(stream, rows, n_in, H, p, p, p, p, p, p, 1 << 40) == -1
    assert b"65535 * 64" in lib.szsnn_last_error(None)
    assert lib.szsnn_adj_input_grad_noalloc(stream, rows, n_in, H, p, p, p, p, p) == -1
    assert b"65535 * 64" in lib.szsnn_last_error(None)
