"""Long-run statistics of the oracle after chaotic divergence (CPU).

The FP32 oracle (the arithmetic the HIP path reproduces bit for bit) and the
FP64 oracle agree spike for spike until a rounding difference flips a first
spike; past that step their firing rates, 10 ms binned population counts and
CV-ISI distributions agree within tests/spike_stats.py's tolerances, while
a planted parameter change (Izhikevich reset increment d x 0.9, IF tau_m x
1.05) falls outside them.  The IF_curr_exp 10k network as configured stays
subthreshold (its neurons never fire), so its FP32 and FP64 rasters are
identical over all 10000 steps.
"""
import dataclasses

import numpy as np
import pytest

from tests import spike_stats as ss


def _run(oracle_mod, net, steps, precision):
    o = oracle_mod.Oracle(net, precision)
    r, _ = o.run(steps, threads=oracle_mod.host_threads() if precision == 32 else 1)
    return r


def _stats(net, raster, skip):
    from szsnn_amd.simulation import raster_to_bool
    return ss.statistics(raster_to_bool(raster, net.n_neurons)[skip:], ss.populations(net), net.meta["dt_ms"])


@pytest.mark.parametrize("name", sorted(ss.CASES))
def test_fp32_fp64_statistics_after_divergence(oracle_mod, name):
    net, steps, skip = ss.build(name)
    r32 = _run(oracle_mod, net, steps, 32)
    r64 = _run(oracle_mod, net, steps, 64)
    div = ss.first_divergence(r32, r64)
    assert div is not None and div < skip + (steps - skip) // 2, div
    s64 = _stats(net, r64, skip)
    ss.assert_statistics_match(_stats(net, r32, skip), s64, ss.TOLS[name])
    # a planted parameter change is caught
    p = net.params.copy()
    if name == "izhikevich2003":
        p[3] *= 0.9   # reset increment d
    else:
        p[2] *= 1.05  # tau_m
    bad = ss.outside(_stats(net, _run(oracle_mod, dataclasses.replace(net, params=p), steps, 32), skip), s64,
                     ss.TOLS[name])
    assert bad, "a 5-10 % parameter change went unnoticed"


def test_if_curr_exp_10k_as_configured_never_diverges(oracle_mod):
    from szsnn_amd.configs import if_curr_exp_10k
    from szsnn_amd.simulation import raster_to_bool
    net = if_curr_exp_10k()
    r32 = _run(oracle_mod, net, 10000, 32)
    r64 = _run(oracle_mod, net, 10000, 64)
    np.testing.assert_array_equal(r32, r64)
    b = raster_to_bool(r32, net.n_neurons + net.n_sources)
    assert not b[:, :net.n_neurons].any() and b[:, net.n_neurons:].sum() > 5000


def test_ks_distance_known_values():
    assert ss.ks_distance([1, 2, 3], [1, 2, 3]) == 0.0
    assert ss.ks_distance([0, 0], [1, 1]) == 1.0
    assert ss.ks_distance([1, 2, 3, 4], [3, 4, 5, 6]) == 0.5
