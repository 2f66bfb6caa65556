"""Batched mode on the GPU: every instance of a batch evolves exactly as the
same network run alone (bit-exact rasters and state), and as the FP32 oracle;
a batch with its gather policy switched on equals the default batch."""
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


@pytest.mark.parametrize("policy,heavy", [(0.0, 0), (0.012, 1)])
def test_batch_with_pulls_and_heavy_pushes_equals_default(ctx, policy, heavy):
    """A batch with its gather policy switched on (every step pulled; or
    the plain runs' fraction with every spiking block posting heavy pushes,
    so steps are deferred, predicted or late) equals the default batch,
    which pushes every step from its update blocks: the pull's target
    blocks span instance boundaries."""
    from szsnn_amd.batch import BatchSimulation, batched_izhikevich
    nets = batched_izhikevich(seed=12, n_instances=5, n_min=300, n_max=2500, fanout=40)
    steps = 250
    print(f"case: batch with policy {policy}, heavy threshold {heavy}", flush=True)
    runs = []
    for on in (False, True):
        bs = BatchSimulation(ctx, nets)
        if on:
            assert bs.gather_policy()["has_pull"]
            bs.set_heavy_push(heavy)
            bs.set_gather_policy(policy)
        bs.set_recording(0, steps)
        bs.run(steps)
        runs.append((bs.raster(), bs.get_vars(), bs.counters()))
        bs.close()
    np.testing.assert_array_equal(runs[1][0], runs[0][0])
    np.testing.assert_array_equal(runs[1][1], runs[0][1])
    for k in ("events", "spikes", "rows", "drive_events"):
        assert runs[1][2][k] == runs[0][2][k], k
    assert runs[0][2]["deferred"] == 0 and runs[1][2]["deferred"] > 0
    assert runs[0][2]["spikes"] > 0
