"""FP64 PyTorch restatement of the feedforward SNN loss (checker for the
HIP adjoint in szsnn_amd/adjoint.py).

TEST INFRASTRUCTURE ONLY: imported by tests/ and bench.py's baseline leg.

Same discretisation as the IF_curr_exp step (IF_curr_exp_update_kernel.cl:
230-297, division as a product with the reciprocal), sign-split synaptic
input with delay 1, readout y_t = alpha y_{t-1} + W2^T (v_t - v_rest), loss
mean((y - target)^2).  With `spikes` given, the spike pattern (and therefore
every reset and refractory mask) is held fixed, so torch autograd yields the
gradient of the piecewise-smooth map the adjoint differentiates.
"""
import torch


def forward(W1, W2, X, target, params, alpha, spikes=None, return_v=False):
    """All tensors float64 (CPU).  params [9, H].  Returns (loss, spikes [T,B,H]
    bool) and, with return_v, the post-update v [T,B,H]."""
    T, B, n_in = X.shape
    H = W1.shape[1]
    v_rest, cm, tau_m, tau_refrac, tau_e, tau_i, i_offset, v_reset, v_thresh = [params[k] for k in range(9)]
    rc, rt = 1.0 / cm, 1.0 / tau_m
    ae, ai = 1.0 - 1.0 / tau_e, 1.0 - 1.0 / tau_i
    W1p = W1.clamp(min=0)
    W1n = W1.clamp(max=0)
    v = v_rest.expand(B, H).clone()
    se = torch.zeros(B, H, dtype=W1.dtype)
    si = torch.zeros(B, H, dtype=W1.dtype)
    fired = torch.full((B, H), 10000.0, dtype=W1.dtype)
    y = torch.zeros(B, W2.shape[1], dtype=W1.dtype)
    loss = torch.zeros((), dtype=W1.dtype)
    out_spikes = []
    vs = []
    for t in range(T):
        x = X[t - 1] if t > 0 else torch.zeros(B, n_in, dtype=W1.dtype)
        se = ae * se + x @ W1p
        si = ai * si + x @ W1n
        I = se + si + i_offset
        refr = fired <= tau_refrac
        fired = fired + 1.0
        vp = torch.where(refr, v, v + (I * rc + (v_rest - v) * rt))
        if spikes is None:
            s = (~refr) & (vp >= v_thresh)
        else:
            s = spikes[t]
        v = torch.where(s, v_reset.expand(B, H), vp)
        fired = torch.where(s, torch.ones_like(fired), fired)
        out_spikes.append(s)
        if return_v:
            vs.append(v.detach().clone())
        y = alpha * y + (v - v_rest) @ W2
        loss = loss + ((y - target[t]) ** 2).sum()
    loss = loss / (T * B * W2.shape[1])
    if return_v:
        return loss, torch.stack(out_spikes), torch.stack(vs)
    return loss, torch.stack(out_spikes)
