"""Network description and construction helpers (host side, numpy).

`Network` is the plain description handed to the HIP library
(szsnn_net_desc in include/szsnn_hip.h).  It mirrors what the reference's
NetworkGraph<NeuronModel, InputModel> holds once built
(szsnn1.1/src/NetworkGraph.hpp:90-135): neurons with per-neuron parameter and
variable arrays, an input population, and synapses (source, target, weight,
API delay).  Parameters are in per-step units because the reference kernels
carry no dt; `if_curr_exp_params` and friends convert PyNN units (ms, nF, mV,
nA, uS) with the step dt.
"""
from dataclasses import dataclass, field
import math

import numpy as np

from ._lib import (MODEL_IF_COND_EXP, MODEL_IF_CURR_EXP, MODEL_IZHIKEVICH, SOURCE_CONST_RATE,
                   SOURCE_NONE, SOURCE_POISSON)

N_PARAMS = {MODEL_IF_CURR_EXP: 9, MODEL_IZHIKEVICH: 4, MODEL_IF_COND_EXP: 11}
N_VARS = {MODEL_IF_CURR_EXP: 4, MODEL_IZHIKEVICH: 2, MODEL_IF_COND_EXP: 4}
MODEL_NAMES = {MODEL_IF_CURR_EXP: "IF_curr_exp", MODEL_IZHIKEVICH: "Izhikevich",
               MODEL_IF_COND_EXP: "IF_cond_exp"}


@dataclass
class Network:
    model: int
    n_neurons: int
    params: np.ndarray            # float32 [n_params, N]
    vars: np.ndarray              # float32 [n_vars, N]
    syn_src: np.ndarray           # uint32 [S]
    syn_tgt: np.ndarray           # uint32 [S]
    syn_w: np.ndarray             # float32 [S]
    syn_delay: np.ndarray         # uint16 [S], API delay k
    max_delay: int
    frac_bits: int
    seed: int = 0
    n_sources: int = 0
    source_model: int = SOURCE_NONE
    src_rate: np.ndarray = None   # float32 [M]
    src_last: np.ndarray = None   # float32 [M]
    src_thresh: np.ndarray = None # uint64 [M]
    noise_amp: np.ndarray = None  # float32 [N]
    drive_cdf: np.ndarray = None  # uint64 [K]
    drive_w: float = 0.0
    queue_capacity: int = 0
    pull_layout: int = 0          # 0: dense-gather layout when possible (2- or 4-byte entries when allowed); 1: 8-byte entries; -1: never
    meta: dict = field(default_factory=dict)

    @property
    def n_syn(self):
        return int(self.syn_src.shape[0])

    def validate(self):
        n = self.n_neurons
        assert self.params.dtype == np.float32 and self.params.shape == (N_PARAMS[self.model], n)
        assert self.vars.dtype == np.float32 and self.vars.shape == (N_VARS[self.model], n)
        for a, dt in ((self.syn_src, np.uint32), (self.syn_tgt, np.uint32), (self.syn_w, np.float32),
                      (self.syn_delay, np.uint16)):
            assert a.dtype == dt and a.ndim == 1 and a.shape[0] == self.n_syn
        return self


# ---------------------------------------------------------------- fixed point
def choose_frac_bits(max_abs_w):
    """Largest F <= 40 with max|w| * 2^F < 2^30 (room for the int32 weight)."""
    if max_abs_w <= 0:
        return 32
    return int(min(40, 30 - math.ceil(math.log2(max_abs_w))))


def quantize_weights(w, frac_bits):
    """Round weights (as float32) to multiples of 2^-F; the result is exactly representable
    in float32 and as an int32 multiple of 2^-F, so the fixed-point sum on the
    device and the float weights seen by any checker are the same numbers."""
    w = np.asarray(w, dtype=np.float32).astype(np.float64)
    q = np.rint(w * 2.0 ** frac_bits) * 2.0 ** -frac_bits
    out = q.astype(np.float32)
    assert np.all(out.astype(np.float64) == q), "quantised weight not representable in float32"
    return out


# ---------------------------------------------------------------- models
def if_curr_exp_params(n, dt=1.0, v_rest=-65.0, cm=1.0, tau_m=20.0, tau_refrac=0.0, tau_syn_E=5.0,
                       tau_syn_I=5.0, i_offset=0.0, v_reset=-65.0, v_thresh=-50.0):
    """IF_curr_exp parameters in per-step units, parameter order of
    szsnn1.1/src/models/IF_curr_exp_NeuronModel.hpp:12-23."""
    vals = [v_rest, cm / dt, tau_m / dt, tau_refrac / dt, tau_syn_E / dt, tau_syn_I / dt, i_offset,
            v_reset, v_thresh]
    return np.stack([np.broadcast_to(np.float32(v), (n,)) if np.ndim(v) == 0
                     else np.asarray(v, np.float32) for v in vals]).astype(np.float32)


def if_cond_exp_params(n, dt=1.0, v_rest=-65.0, cm=1.0, tau_m=20.0, tau_refrac=0.0, tau_syn_E=5.0,
                       tau_syn_I=5.0, i_offset=0.0, v_reset=-65.0, v_thresh=-50.0, e_rev_E=0.0,
                       e_rev_I=-70.0):
    base = if_curr_exp_params(n, dt, v_rest, cm, tau_m, tau_refrac, tau_syn_E, tau_syn_I, i_offset,
                              v_reset, v_thresh)
    extra = np.stack([np.full(n, e_rev_E, np.float32), np.full(n, e_rev_I, np.float32)])
    return np.concatenate([base, extra]).astype(np.float32)


def if_vars(n, v0=-65.0):
    """[v, ie|ge, ii|gi, fired]; fired starts at 10000 (IF_curr_exp_NeuronModel.hpp:47-51)."""
    v = np.broadcast_to(np.asarray(v0, np.float32), (n,))
    return np.stack([v, np.zeros(n, np.float32), np.zeros(n, np.float32),
                     np.full(n, 10000.0, np.float32)]).astype(np.float32)


def izhikevich_params(a, b, c, d):
    return np.stack([np.asarray(x, np.float32) for x in (a, b, c, d)]).astype(np.float32)


def poisson_thresholds(p):
    """Integer thresholds: a source fires when its 64-bit draw is < p * 2^64."""
    p = np.asarray(p, dtype=np.float64)
    out = np.empty(p.shape, np.uint64)
    for i, x in np.ndenumerate(p):
        out[i] = min(int(x * 2.0 ** 64), 2 ** 64 - 1)
    return out


def binomial_cdf_thresholds(n, p):
    """Thresholds t_j = 2^64 * P(X <= j), X ~ Binomial(n, p), for j while
    P(X > j) >= 2^-64.  events = #{j : u >= t_j} for a uniform 64-bit u."""
    from scipy.stats import binom
    out = []
    j = 0
    while True:
        sf = float(binom.sf(j, n, p))  # P(X > j), accurate in the tail
        tail = int(round(sf * 2.0 ** 64))
        if tail <= 0 or j >= n:
            break
        out.append(2 ** 64 - tail)
        j += 1
    return np.array(out, dtype=np.uint64)


# ---------------------------------------------------------------- sampling
def distinct_choice(rng, n_rows, k, pool, exclude=None, dtype=np.int64):
    """(n_rows, k) array; row r holds k distinct values from range(pool),
    none equal to exclude[r] (if given)."""
    m = pool - (0 if exclude is None else 1)
    if k > m:
        raise ValueError("cannot draw k distinct values from the pool")
    x = rng.integers(0, m, size=(n_rows, k), dtype=dtype)
    while True:
        x.sort(axis=1)
        dup = np.zeros(x.shape, dtype=bool)
        dup[:, 1:] = x[:, 1:] == x[:, :-1]
        nd = int(dup.sum())
        if nd == 0:
            break
        x[dup] = rng.integers(0, m, size=nd, dtype=dtype)
    if exclude is not None:
        ex = np.asarray(exclude, dtype=dtype)[:, None]
        x = x + (x >= ex)
    return x


def concat_synapses(parts):
    src = np.concatenate([p[0] for p in parts]).astype(np.uint32)
    tgt = np.concatenate([p[1] for p in parts]).astype(np.uint32)
    w = np.concatenate([p[2] for p in parts]).astype(np.float32)
    d = np.concatenate([p[3] for p in parts]).astype(np.uint16)
    return src, tgt, w, d
