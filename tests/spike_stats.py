"""Long-run spike statistics for comparing runs after chaotic divergence.

Two runs of one network in different arithmetic (FP32 and FP64) agree spike
for spike only until rounding differences flip a first spike; after that
they are compared as statistics, the way the reference compared its rasters
with other simulators (population spike counts summed in 10 ms bins,
pyszsnn/src/pyszsnn/print_spikesv1.py:68-69):

  * mean firing rate per population (Hz);
  * population spike counts in 10 ms bins: mean, and the two-sample
    Kolmogorov-Smirnov distance of their distributions;
  * CV of the inter-spike intervals of every neuron with at least 4 spikes:
    mean, and the KS distance of the distributions.

Tolerances are per network (TOLS), a few times the FP32/FP64 differences
measured on it, and tight enough that a 10 % change of the Izhikevich
reset increment d, or a 5 % change of the IF membrane time constant, falls
outside them (tests/test_longrun_stats.py plants both).
"""
import numpy as np

BIN_MS = 10.0
# Keys: rate_rel (per-population mean rate, relative), binned_mean_rel (mean
# population count per 10 ms bin, relative), binned_ks (KS distance of the
# binned-count distributions), cv_mean_abs (mean CV of the inter-spike
# intervals), cv_ks (KS distance of the CV-ISI distributions).
TOLS = {
    # Izhikevich 2003 net, 1000 neurons, steps 300..1000 (FP32/FP64 measured:
    # 0.0094, 0.0016, 0.057, 0.0004, 0.0096)
    "izhikevich2003": {"rate_rel": 0.03, "binned_mean_rel": 0.03, "binned_ks": 0.2, "cv_mean_abs": 0.01,
                       "cv_ks": 0.05},
    # IF_curr_exp 10k net with 100 Hz / 0.7 nA Poisson sources, steps
    # 1000..10000 (measured: 0.00056, 0.00042, 0.067, 0.0006, 0.0046)
    "if_curr_exp_10k_driven": {"rate_rel": 0.005, "binned_mean_rel": 0.005, "binned_ks": 0.15,
                               "cv_mean_abs": 0.003, "cv_ks": 0.02},
}


def ks_distance(a, b):
    """Two-sample Kolmogorov-Smirnov statistic sup |F_a - F_b|."""
    a = np.sort(np.asarray(a, np.float64))
    b = np.sort(np.asarray(b, np.float64))
    x = np.concatenate([a, b])
    fa = np.searchsorted(a, x, side="right") / a.size
    fb = np.searchsorted(b, x, side="right") / b.size
    return float(np.abs(fa - fb).max())


def statistics(spikes, populations, dt_ms):
    """spikes: bool [steps, neurons]; populations: {name: (first, end)}."""
    steps = spikes.shape[0]
    dur_s = steps * dt_ms * 1e-3
    rates = {k: float(spikes[:, a:b].sum()) / (b - a) / dur_s for k, (a, b) in populations.items()}
    per = int(round(BIN_MS / dt_ms))
    nb = steps // per
    binned = spikes[:nb * per].sum(axis=1).reshape(nb, per).sum(axis=1).astype(np.float64)
    cvs = []
    for i in range(spikes.shape[1]):
        t = np.flatnonzero(spikes[:, i])
        if t.size >= 4:
            isi = np.diff(t).astype(np.float64)
            cvs.append(isi.std() / isi.mean())
    return {"rates": rates, "binned": binned, "cv": np.array(cvs)}


def compare(a, b):
    """Differences between two statistics() results, in the units of TOLS."""
    out = {}
    out["rate_rel"] = max(abs(a["rates"][k] - b["rates"][k]) / max(b["rates"][k], 1e-12) for k in b["rates"])
    out["binned_mean_rel"] = abs(a["binned"].mean() - b["binned"].mean()) / max(b["binned"].mean(), 1e-12)
    out["binned_ks"] = ks_distance(a["binned"], b["binned"])
    out["cv_mean_abs"] = abs(a["cv"].mean() - b["cv"].mean())
    out["cv_ks"] = ks_distance(a["cv"], b["cv"])
    return out


# The long-run cases: (TOLS key, builder arguments, steps, first step compared)
CASES = {
    "izhikevich2003": ({}, 1000, 300),
    "if_curr_exp_10k_driven": ({"poisson_hz": 100.0, "w_poisson": 0.7}, 10000, 1000),
}


def build(name):
    from szsnn_amd import configs
    kwargs, steps, skip = CASES[name]
    net = configs.izhikevich2003(**kwargs) if name == "izhikevich2003" else configs.if_curr_exp_10k(**kwargs)
    return net, steps, skip


def populations(net):
    ne = net.meta["n_exc"]
    return {"exc": (0, ne), "inh": (ne, net.n_neurons)}


def first_divergence(ra, rb):
    """First step at which two rasters differ (None if identical)."""
    d = np.flatnonzero((np.asarray(ra) != np.asarray(rb)).any(axis=1))
    return int(d[0]) if d.size else None


def outside(sa, sb, tol):
    """{key: (difference, tolerance)} of the statistics outside tolerance."""
    diff = compare(sa, sb)
    return {k: (float(v), tol[k]) for k, v in diff.items() if not v <= tol[k]}


def assert_statistics_match(sa, sb, tol):
    bad = outside(sa, sb, tol)
    assert not bad, f"statistics outside tolerance (value, tolerance): {bad}"
