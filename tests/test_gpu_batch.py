"""Batched mode on the GPU: every instance of a batch evolves exactly as the
same network run alone (bit-exact rasters and state), and as the FP32 oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from szsnn_amd.simulation import Context
    return Context(0)


def _plain(ctx, net, steps):
    from szsnn_amd.simulation import Simulation, raster_to_bool
    s = Simulation(ctx, net)
    s.set_recording(0, steps)
    s.run(steps)
    return raster_to_bool(s.raster(), net.n_neurons + net.n_sources), s.get_vars()


def test_batched_izhikevich_equals_plain_runs(ctx):
    from szsnn_amd.batch import BatchSimulation, batched_izhikevich
    nets = batched_izhikevich(seed=11, n_instances=7, n_min=100, n_max=3000, fanout=40)
    steps = 300
    print("case: batch of", [n.n_neurons for n in nets], flush=True)
    bs = BatchSimulation(ctx, nets)
    # a batch pushes every step from its update blocks (no pull, no heavy
    # pushes: graph steps without gather launches); the plain runs keep the
    # default policy, which pulls and posts heavy pushes
    assert bs.heavy_push() == 0 and bs.gather_policy()["dense_fraction"] > 1.0
    bs.set_recording(0, steps)
    bs.run(steps)
    r = bs.raster()
    v = bs.get_vars()
    for b, net in enumerate(nets):
        pr, pv = _plain(ctx, net, steps)
        np.testing.assert_array_equal(bs.instance_raster(b, r), pr)
        np.testing.assert_array_equal(bs.instance_vars(b, v), pv)
    from oracle.pyoracle import Oracle
    from szsnn_amd.simulation import raster_to_bool
    o = Oracle(nets[3], 32)
    ro, _ = o.run(steps)
    np.testing.assert_array_equal(bs.instance_raster(3, r), raster_to_bool(ro, nets[3].n_neurons))


def test_batched_with_sources_equals_plain(ctx):
    from szsnn_amd.batch import BatchSimulation
    from szsnn_amd.configs import if_curr_exp_10k
    nets = [if_curr_exp_10k(seed=40 + i, n_exc=400 + 100 * i, n_inh=100 + 30 * i, n_poisson=50 + 20 * i,
                            poisson_hz=200.0, w_poisson=2.0) for i in range(4)]
    steps = 400
    bs = BatchSimulation(ctx, nets)
    bs.set_recording(0, steps)
    bs.run(steps)
    r = bs.raster()
    total = 0
    for b, net in enumerate(nets):
        pr, pv = _plain(ctx, net, steps)
        ir = bs.instance_raster(b, r)
        np.testing.assert_array_equal(ir, pr)
        total += int(ir.sum())
    assert total == bs.counters()["spikes"]
