"""Synthetic networks for the benchmark configurations (seeded, deterministic).

Each builder takes the sizes as arguments so tests can build scaled-down
versions of the same construction.  "Delay of d steps" means API delay d-1
(a spike of step c reaches the update of step c+d; NetworkGraph.ipp:416 and
scatter/gather slot arithmetic).  Default seed 20261203.
"""
import numpy as np

from ._lib import MODEL_IF_COND_EXP, MODEL_IF_CURR_EXP, MODEL_IZHIKEVICH, SOURCE_POISSON
from .network import (Network, binomial_cdf_thresholds, choose_frac_bits, concat_synapses,
                      distinct_choice, if_cond_exp_params, if_curr_exp_params, if_vars,
                      izhikevich_params, poisson_thresholds, quantize_weights)

DEFAULT_SEED = 20261203


def izhikevich2003(seed=DEFAULT_SEED, n_exc=800, n_inh=200, fanout=100, max_delay_steps=20,
                   noise_exc=5.0, noise_inh=2.0):
    """Izhikevich (2003) network: per-neuron heterogeneous a,b,c,d from r~U[0,1),
    `fanout` distinct random targets per neuron (no self), delays U{1..max}
    steps, weights exc 0.5*U / inh -U, Gaussian current noise_exc / noise_inh
    times N(0,1) per step, dt = 1 ms."""
    rng = np.random.default_rng(seed)
    n = n_exc + n_inh
    re = rng.random(n_exc)
    ri = rng.random(n_inh)
    a = np.concatenate([np.full(n_exc, 0.02), 0.02 + 0.08 * ri])
    b = np.concatenate([np.full(n_exc, 0.2), 0.25 - 0.05 * ri])
    c = np.concatenate([-65.0 + 15.0 * re ** 2, np.full(n_inh, -65.0)])
    d = np.concatenate([8.0 - 6.0 * re ** 2, np.full(n_inh, 2.0)])
    params = izhikevich_params(a, b, c, d)
    v0 = np.full(n, -65.0, np.float32)
    u0 = (params[1] * v0).astype(np.float32)
    vars0 = np.stack([v0, u0]).astype(np.float32)
    k = min(fanout, n - 1)
    tgt = distinct_choice(rng, n, k, n, exclude=np.arange(n)).ravel()
    src = np.repeat(np.arange(n), k)
    delay = rng.integers(1, max_delay_steps + 1, size=src.shape[0]) - 1
    w = rng.random(src.shape[0])
    w = np.where(src < n_exc, 0.5 * w, -w)
    frac = choose_frac_bits(1.0)
    noise = np.concatenate([np.full(n_exc, noise_exc), np.full(n_inh, noise_inh)]).astype(np.float32)
    return Network(model=MODEL_IZHIKEVICH, n_neurons=n, params=params, vars=vars0,
                   syn_src=src.astype(np.uint32), syn_tgt=tgt.astype(np.uint32),
                   syn_w=quantize_weights(w, frac), syn_delay=delay.astype(np.uint16),
                   max_delay=max_delay_steps, frac_bits=frac, seed=seed, noise_amp=noise,
                   meta={"config": "izhikevich2003", "dt_ms": 1.0, "n_exc": n_exc, "n_inh": n_inh}).validate()


def _fixed_indegree(rng, n_post, pre_lo, pre_hi, k, exclude_self):
    """k distinct presynaptic ids in [pre_lo, pre_hi) for each target; a
    target inside the range never draws itself.  Draws k+1 distinct ids, then
    drops the target itself if present, else one uniformly chosen id: both
    leave a uniform k-subset of the allowed ids."""
    pool = pre_hi - pre_lo
    tgt = np.arange(n_post)
    if not exclude_self:
        x = distinct_choice(rng, n_post, k, pool)
        return (x + pre_lo).ravel(), np.repeat(tgt, k)
    x = distinct_choice(rng, n_post, k + 1, pool)
    own = np.where((tgt >= pre_lo) & (tgt < pre_hi), tgt - pre_lo, -1)
    hit = x == own[:, None]
    has = hit.any(axis=1)
    drop = np.where(has, hit.argmax(axis=1), rng.integers(0, k + 1, n_post))
    keep = np.ones(x.shape, dtype=bool)
    keep[np.arange(n_post), drop] = False
    x = x[keep].reshape(n_post, k)
    return (x + pre_lo).ravel(), np.repeat(tgt, k)


def if_curr_exp_10k(seed=DEFAULT_SEED, n_exc=8000, n_inh=2000, k_exc=80, k_inh=20, w_exc=0.05,
                    w_inh=-0.25, max_delay_steps=16, n_poisson=1000, poisson_hz=10.0,
                    poisson_fanout=50, w_poisson=0.1, poisson_delay_steps=1, dt=0.1):
    """IF_curr_exp recurrent network: fixed in-degree k_exc + k_inh (distinct,
    no self), delays U{1..max} steps, Poisson sources each projecting to
    `poisson_fanout` distinct neurons.  PyNN-style parameters at dt (ms);
    initial v ~ U[v_reset, v_thresh)."""
    rng = np.random.default_rng(seed)
    n = n_exc + n_inh
    params = if_curr_exp_params(n, dt=dt, v_rest=-65.0, cm=1.0, tau_m=20.0, tau_refrac=2.0,
                                tau_syn_E=5.0, tau_syn_I=5.0, v_reset=-65.0, v_thresh=-50.0)
    vars0 = if_vars(n, rng.uniform(-65.0, -50.0, n).astype(np.float32))
    se, te = _fixed_indegree(rng, n, 0, n_exc, min(k_exc, n_exc - 1), True)
    si, ti = _fixed_indegree(rng, n, n_exc, n, min(k_inh, n_inh - 1), True)
    de = rng.integers(1, max_delay_steps + 1, size=se.shape[0]) - 1
    di = rng.integers(1, max_delay_steps + 1, size=si.shape[0]) - 1
    parts = [(se, te, np.full(se.shape[0], w_exc), de), (si, ti, np.full(si.shape[0], w_inh), di)]
    if n_poisson:
        pt = distinct_choice(rng, n_poisson, min(poisson_fanout, n), n).ravel()
        ps = np.repeat(np.arange(n, n + n_poisson), min(poisson_fanout, n))
        parts.append((ps, pt, np.full(ps.shape[0], w_poisson),
                      np.full(ps.shape[0], poisson_delay_steps - 1)))
    src, tgt, w, dl = concat_synapses(parts)
    frac = choose_frac_bits(max(abs(w_exc), abs(w_inh), abs(w_poisson)))
    p = poisson_hz * dt * 1e-3
    return Network(model=MODEL_IF_CURR_EXP, n_neurons=n, params=params, vars=vars0, syn_src=src,
                   syn_tgt=tgt, syn_w=quantize_weights(w, frac), syn_delay=dl,
                   max_delay=max(max_delay_steps, poisson_delay_steps), frac_bits=frac, seed=seed,
                   n_sources=n_poisson, source_model=SOURCE_POISSON,
                   src_thresh=poisson_thresholds(np.full(n_poisson, p)),
                   meta={"config": "if_curr_exp_10k", "dt_ms": dt, "n_exc": n_exc, "n_inh": n_inh}).validate()


def if_cond_exp_100k(seed=DEFAULT_SEED, n_exc=80000, n_inh=20000, k_exc=800, k_inh=200,
                     g_exc=0.004, g_inh=0.051, max_delay_steps=32, drive_n=100, drive_hz=5.0,
                     dt=0.1):
    """IF_cond_exp network: fixed in-degree k_exc + k_inh (distinct, no self),
    conductance weights g_exc / -g_inh (inhibitory stored negative), delays
    U{1..max} steps, per-neuron external drive Binomial(drive_n, drive_hz*dt)
    events per step of weight g_exc into g_e.  e_rev_E = 0, e_rev_I = -80 mV,
    tau_syn_E = 5, tau_syn_I = 10 ms, other parameters as the IF_curr_exp net;
    initial v ~ U[v_reset, v_thresh)."""
    rng = np.random.default_rng(seed)
    n = n_exc + n_inh
    params = if_cond_exp_params(n, dt=dt, v_rest=-65.0, cm=1.0, tau_m=20.0, tau_refrac=2.0,
                                tau_syn_E=5.0, tau_syn_I=10.0, v_reset=-65.0, v_thresh=-50.0,
                                e_rev_E=0.0, e_rev_I=-80.0)
    vars0 = if_vars(n, rng.uniform(-65.0, -50.0, n).astype(np.float32))
    se, te = _fixed_indegree(rng, n, 0, n_exc, min(k_exc, n_exc - 1), True)
    si, ti = _fixed_indegree(rng, n, n_exc, n, min(k_inh, n_inh - 1), True)
    de = rng.integers(1, max_delay_steps + 1, size=se.shape[0], dtype=np.int32) - 1
    di = rng.integers(1, max_delay_steps + 1, size=si.shape[0], dtype=np.int32) - 1
    src, tgt, w, dl = concat_synapses([(se, te, np.full(se.shape[0], g_exc, np.float32), de),
                                       (si, ti, np.full(si.shape[0], -g_inh, np.float32), di)])
    del se, te, si, ti, de, di
    frac = choose_frac_bits(max(g_exc, g_inh))
    cdf = binomial_cdf_thresholds(drive_n, drive_hz * dt * 1e-3) if drive_n else None
    return Network(model=MODEL_IF_COND_EXP, n_neurons=n, params=params, vars=vars0, syn_src=src,
                   syn_tgt=tgt, syn_w=quantize_weights(w, frac), syn_delay=dl,
                   max_delay=max_delay_steps, frac_bits=frac, seed=seed, drive_cdf=cdf,
                   drive_w=float(quantize_weights([g_exc], frac)[0]),
                   meta={"config": "if_cond_exp_100k", "dt_ms": dt, "n_exc": n_exc, "n_inh": n_inh}).validate()
