"""Differentiable IF_curr_exp feedforward network: forward and discrete
adjoint on the GPU, exposed as a torch.autograd.Function.

Network (one trial b, step t; dt = 1 ms):
  x_t      Poisson input spikes [n_in], fixed data
  hidden   IF_curr_exp (szsnn1.1/src/cl/IF_curr_exp_update_kernel.cl:230-297 in
           this library's FP32 form), synaptic input W1^T x_{t-1} (delay 1),
           sign-split into the E (w >= 0) and I (w < 0) currents
  readout  y_t = alpha y_{t-1} + W2^T h_t, with h_t = v_t - v_rest: the
           readout is driven by the hidden depolarisation, so it depends
           continuously on W1 while the spike pattern is held fixed (a readout
           driven only by spikes would have dL/dW1 = 0 under a fixed pattern)
  loss     mean over (t, b, k) of (y_t - target_t)^2

The forward computes the input currents of all T steps with one HIP kernel
that adds only the spiking inputs' rows of W1 (split by sign), runs the
recurrence segment by segment (the hidden state is
stored every K steps, with the spike bits and h of every step), and forms the
readout drive in one GEMM.  The backward runs the hand-written adjoint
kernel backwards through each segment from its checkpoint, then forms dW1
with a HIP kernel that walks, per input, only the steps where it spiked
(about 2 % of the (step, trial) rows), and dW2 in one GEMM (about 0.4 GB
of per-step currents, h and adjoints stay resident for T = 200, B = 64,
H = 2000).  The readout GEMMs go to the vendor BLAS through torch.matmul;
the recurrences and the input sums are the HIP kernels of szsnn_adjoint.hip.
"""
import ctypes
import math

import numpy as np

from . import _lib

IF_PARAM_ORDER = ("v_rest", "cm", "tau_m", "tau_refrac", "tau_syn_E", "tau_syn_I", "i_offset", "v_reset",
                  "v_thresh")


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream(torch):
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def hidden_params(n_hidden, dt=1.0, v_rest=-65.0, cm=1.0, tau_m=20.0, tau_refrac=2.0, tau_syn_E=5.0,
                  tau_syn_I=5.0, i_offset=0.0, v_reset=-65.0, v_thresh=-50.0):
    """[9, H] per-step IF_curr_exp parameters (the IF_curr_exp 10k network's)."""
    from .network import if_curr_exp_params
    return if_curr_exp_params(n_hidden, dt, v_rest, cm, tau_m, tau_refrac, tau_syn_E, tau_syn_I, i_offset,
                              v_reset, v_thresh)


def make_problem(seed=20261203, batch=64, n_in=1000, n_hidden=2000, n_out=100, steps=200, dt=1.0,
                 rate_lo=2.0, rate_hi=40.0, w1_mean=0.0, w1_std=0.02, w2_std=None, tau_out=20.0,
                 n_sines=3, f_max_hz=10.0):
    """Seeded problem: input spikes X [T, B, n_in] (Poisson, per-input rates
    U[rate_lo, rate_hi] Hz), W1 ~ N(w1_mean, w1_std) [n_in, H], W2 ~
    N(0, w2_std) [H, n_out], targets [T, B, n_out] (sums of n_sines random
    sinusoids below f_max_hz).  numpy float32 arrays."""
    rng = np.random.default_rng(seed)
    rates = rng.uniform(rate_lo, rate_hi, n_in)
    p = rates * dt * 1e-3
    X = (rng.random((steps, batch, n_in)) < p).astype(np.float32)
    W1 = rng.normal(w1_mean, w1_std, (n_in, n_hidden)).astype(np.float32)
    if w2_std is None:
        w2_std = 1.0 / math.sqrt(n_hidden)
    W2 = rng.normal(0.0, w2_std, (n_hidden, n_out)).astype(np.float32)
    t = np.arange(steps) * dt * 1e-3
    amp = rng.normal(0, 1, (n_sines, batch, n_out))
    freq = rng.uniform(0.5, f_max_hz, (n_sines, batch, n_out))
    ph = rng.uniform(0, 2 * np.pi, (n_sines, batch, n_out))
    target = np.zeros((steps, batch, n_out))
    for s in range(n_sines):
        target += amp[s] * np.sin(2 * np.pi * freq[s] * t[:, None, None] + ph[s])
    return {"X": X, "W1": W1, "W2": W2, "target": target.astype(np.float32),
            "params": hidden_params(n_hidden, dt=dt), "alpha": float(np.float32(1.0 - dt / tau_out))}


def initial_state(torch, B, H, params, device):
    st = torch.zeros(4, B * H, dtype=torch.float32, device=device)
    st[0] = params[0].repeat(B)          # v = v_rest
    st[3] = 10000.0                      # fired counter: long since the last spike
    return st


def _check(rc, what):
    if rc != _lib.SZSNN_OK:
        raise _lib.SzsnnError(rc, what + " failed")


class SNNLoss:
    """Builds the autograd Function lazily (torch imported on use)."""

    _fn = None

    @classmethod
    def function(cls):
        if cls._fn is None:
            cls._fn = _make_function()
        return cls._fn


def _make_function():
    import torch

    lib = _lib.load()

    class _F(torch.autograd.Function):
        @staticmethod
        def forward(ctx, W1, W2, X, target, params, alpha, K):
            T, B, n_in = X.shape
            H = W1.shape[1]
            R = W2.shape[1]
            dev = W1.device
            Xprev = torch.cat([torch.zeros(1, B, n_in, device=dev, dtype=X.dtype), X[:-1]], 0).contiguous()
            xs = Xprev.reshape(T * B, n_in).to(torch.float32).contiguous()
            # the input currents of every step (only the spiking inputs are
            # added), kept for the backward
            AE = torch.empty(T, B * H, dtype=torch.float32, device=dev)
            AI = torch.empty(T, B * H, dtype=torch.float32, device=dev)
            W1c = W1.contiguous()
            st = _stream(torch)
            _check(lib.szsnn_adj_input_currents(st, T * B, n_in, H, _ptr(xs), _ptr(W1c), _ptr(AE), _ptr(AI)),
                   "adj_input_currents")
            words = (B * H + 63) // 64
            bits = torch.zeros(T, words, dtype=torch.int64, device=dev)
            h = torch.empty(T, B * H, dtype=torch.float32, device=dev)
            state = initial_state(torch, B, H, params, dev)
            ckpts = []
            for s0 in range(0, T, K):
                k = min(K, T - s0)
                ckpts.append(state.clone())
                _check(lib.szsnn_adj_forward_segment(st, B, H, k, _ptr(params), _ptr(state), _ptr(AE[s0]),
                                                     _ptr(AI[s0]), _ptr(h[s0]), _ptr(bits[s0])),
                       "adj_forward_segment")
            drive = (h.reshape(T * B, H) @ W2).reshape(T, B, R).contiguous()
            y = torch.empty(T, B, R, dtype=torch.float32, device=dev)
            _check(lib.szsnn_readout_forward(st, B * R, T, alpha, None, _ptr(drive), _ptr(y)), "readout_forward")
            loss = ((y - target) ** 2).mean()
            ctx.save_for_backward(W1, W2, xs, target, params, y, bits, h, *ckpts)
            ctx.alpha = alpha
            ctx.K = K
            ctx.mark_non_differentiable(y, bits)
            return loss, y, bits

        @staticmethod
        def backward(ctx, gloss, gy, gbits):
            W1, W2, xs, target, params, y, bits, h, *ckpts = ctx.saved_tensors
            alpha, K = ctx.alpha, ctx.K
            T, R = y.shape[0], y.shape[2]
            B = y.shape[1]
            H = W1.shape[1]
            dev = W1.device
            st = _stream(torch)
            lam = torch.empty(T, B, R, dtype=torch.float32, device=dev)
            scale = float(gloss) * 2.0 / (T * B * R)
            _check(lib.szsnn_readout_backward(st, B * R, T, alpha, scale, _ptr(y), _ptr(target), _ptr(lam)),
                   "readout_backward")
            lam2 = lam.reshape(T * B, R)
            dW2 = h.reshape(T * B, H).t() @ lam2
            g = (lam2 @ W2.t()).reshape(T, B * H).contiguous()
            carry = torch.zeros(3, B * H, dtype=torch.float32, device=dev)
            LE = torch.empty(T, B * H, dtype=torch.float32, device=dev)
            LI = torch.empty(T, B * H, dtype=torch.float32, device=dev)
            starts = list(range(0, T, K))
            for i in reversed(range(len(starts))):
                s0 = starts[i]
                k = min(K, T - s0)
                _check(lib.szsnn_adj_backward_segment(st, B, H, k, _ptr(params), _ptr(ckpts[i]), _ptr(bits[s0]),
                                                      _ptr(g[s0]), _ptr(carry), _ptr(LE[s0]), _ptr(LI[s0])),
                       "adj_backward_segment")
            # dW1 over the spiking inputs only (a few percent of xs), each
            # weight taking the adjoint of its sign's current
            n_in = xs.shape[1]
            wbytes = lib.szsnn_adj_input_grad_workspace(T * B, n_in)
            ws = torch.empty((wbytes + 7) // 8, dtype=torch.int64, device=dev)
            dW1 = torch.empty(n_in, H, dtype=torch.float32, device=dev)
            _check(lib.szsnn_adj_input_grad(st, T * B, n_in, H, _ptr(xs), _ptr(W1.contiguous()), _ptr(LE), _ptr(LI),
                                            _ptr(dW1), _ptr(ws), wbytes), "adj_input_grad")
            return dW1, dW2, None, None, None, None, None

    return _F


def snn_loss(W1, W2, X, target, params, alpha, K=20):
    """Loss of the feedforward SNN; differentiable w.r.t. W1 and W2 through
    the HIP adjoint.  Returns (loss, readout y [T,B,R], spike bits [T, words])."""
    if K < 1 or K > 64:
        raise ValueError("checkpoint interval K must be in 1..64")
    return SNNLoss.function().apply(W1, W2, X, target, params, float(alpha), int(K))
