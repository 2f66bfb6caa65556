"""Long-run statistics of the HIP path against the FP64 oracle (GPU).

Past the first step where the HIP (FP32) raster and the FP64 oracle's raster
differ, per-population firing rates, 10 ms binned population counts and the
CV-ISI distribution agree within tests/spike_stats.py's per-network
tolerances (the reference's own comparison method: spike counts in 10 ms
bins against another simulator, pyszsnn/src/pyszsnn/print_spikesv1.py:68-69).
The IF_curr_exp 10k network as configured stays subthreshold and never
diverges: its HIP raster equals the FP64 oracle's over all 10000 steps.
"""
import numpy as np
import pytest

from tests import spike_stats as ss

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from szsnn_amd.simulation import Context
    return Context(0)


def _hip(ctx, net, steps):
    from szsnn_amd.simulation import Simulation
    s = Simulation(ctx, net)
    s.set_recording(0, steps)
    s.run(steps)
    r = s.raster()
    s.close()
    return r


def _fp64(net, steps):
    from oracle.pyoracle import Oracle
    r, _ = Oracle(net, 64).run(steps)
    return r


@pytest.mark.parametrize("name", sorted(ss.CASES))
def test_statistics_against_fp64_after_divergence(ctx, name):
    from szsnn_amd.simulation import raster_to_bool
    net, steps, skip = ss.build(name)
    print(f"case: long-run statistics {name}, {steps} steps from step {skip}", flush=True)
    rg = _hip(ctx, net, steps)
    r64 = _fp64(net, steps)
    div = ss.first_divergence(rg, r64)
    assert div is not None and div < skip + (steps - skip) // 2, div
    pops = ss.populations(net)
    dt = net.meta["dt_ms"]
    sg = ss.statistics(raster_to_bool(rg, net.n_neurons)[skip:], pops, dt)
    s64 = ss.statistics(raster_to_bool(r64, net.n_neurons)[skip:], pops, dt)
    ss.assert_statistics_match(sg, s64, ss.TOLS[name])
    assert sg["cv"].size > net.n_neurons // 2


def test_if_curr_exp_10k_as_configured_equals_fp64(ctx):
    from szsnn_amd.configs import if_curr_exp_10k
    net = if_curr_exp_10k()
    print("case: if_curr_exp_10k as configured, 10000 steps vs FP64", flush=True)
    np.testing.assert_array_equal(_hip(ctx, net, 10000), _fp64(net, 10000))
