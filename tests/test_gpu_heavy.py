"""Heavy pushes: an update block whose spiking ids carry many synapse entries
hands their ranges to the next step's gather launch, which adds them with its
whole grid.  The sums are the same integer additions in another order, so
every result is bit-identical to the block pushing them itself, and to the
FP32 oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from szsnn_amd.simulation import Context
    return Context(0)


def _run(ctx, net, steps, policy, heavy, volley=0, parts=None, persistent_from=None):
    """Runs `steps` steps in the chunks `parts` (default one chunk); chunk
    index persistent_from onwards runs in the persistent mode."""
    from szsnn_amd._lib import MODE_PERSISTENT
    from szsnn_amd.simulation import Simulation
    s = Simulation(ctx, net)
    assert s.gather_policy()["has_pull"]
    s.set_heavy_push(heavy)
    assert s.heavy_push() == heavy
    s.set_gather_policy(policy)
    s.set_volley(volley)
    s.set_recording(0, steps)
    parts = parts or [steps]
    assert sum(parts) == steps
    for i, p in enumerate(parts):
        if persistent_from is not None and i == persistent_from:
            s.set_mode(MODE_PERSISTENT, 256, 0)
        s.run(p)
    return s.raster(), s.get_vars(), s.counters()


def _same(a, b):
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])
    assert a[2]["events"] == b[2]["events"] and a[2]["spikes"] == b[2]["spikes"]
    assert a[2]["drive_events"] == b[2]["drive_events"]


@pytest.mark.parametrize("which", ["izh", "curr"])
def test_every_block_posting_matches_oracle(ctx, which):
    """Threshold 1: every update block with a spiking id posts; pushing every
    step, the rasters, state and event counts equal the FP32 oracle's."""
    from oracle.pyoracle import Oracle
    from szsnn_amd import configs
    if which == "izh":
        net, steps = configs.izhikevich2003(), 300
    else:
        net, steps = configs.if_curr_exp_10k(n_exc=2000, n_inh=500, n_poisson=300, poisson_hz=100.0,
                                             w_poisson=1.5), 400
    print(f"case: heavy threshold 1, push only, {which}", flush=True)
    r, v, c = _run(ctx, net, steps, 2.0, 1)
    o = Oracle(net, 32)
    ro, _ = o.run(steps)
    bad = np.flatnonzero((r != ro).any(axis=1))
    assert bad.size == 0, f"raster differs first at step {bad[0]}"
    np.testing.assert_array_equal(v, o.vars().astype(np.float32))
    assert c["events"] == o.events > 0
    assert c["drive_events"] == o.drive_events


def test_concentrated_burst_equals_block_push_and_pull(ctx):
    """A sustained volley of 250 contiguous forced ids (one update block, 62k
    synapse entries per step): heavy pushes at the default threshold, the
    block pushing them itself, and the pull give identical results."""
    from szsnn_amd import configs
    from szsnn_amd.simulation import Simulation
    net = configs.if_cond_exp_100k(n_exc=4000, n_inh=1000, k_exc=200, k_inh=50)
    steps, volley = 60, 250
    default = Simulation(ctx, net).heavy_push()
    assert default > 0
    print("case: concentrated burst, default threshold vs none vs pull", flush=True)
    heavy = _run(ctx, net, steps, 2.0, default, volley)
    block = _run(ctx, net, steps, 2.0, 0, volley)
    pull = _run(ctx, net, steps, 0.0, default, volley)
    _same(heavy, block)
    _same(heavy, pull)
    assert heavy[2]["events"] > steps * volley * 200


@pytest.mark.parametrize("parts,persistent_from", [([37, 63], 1), ([37, 41, 22], None)])
def test_posts_pending_across_launch_changes(ctx, parts, persistent_from):
    """The ranges posted by the last step of a run are added before the next
    one: by its first gather launch, or by the launch before a persistent
    loop; a threshold change in between adds them first."""
    from szsnn_amd import configs
    from szsnn_amd.simulation import Simulation
    net = configs.izhikevich2003()
    steps = sum(parts)
    print(f"case: pending heavy posts, parts {parts}, persistent from {persistent_from}", flush=True)
    want = _run(ctx, net, steps, 2.0, 0)
    if persistent_from is not None:
        got = _run(ctx, net, steps, 2.0, 1, parts=parts, persistent_from=persistent_from)
    else:
        s = Simulation(ctx, net)
        s.set_heavy_push(1)
        s.set_gather_policy(2.0)
        s.set_recording(0, steps)
        s.run(parts[0])
        s.set_heavy_push(0)      # the posts of step parts[0] - 1 are pushed here
        s.run(parts[1])
        s.set_heavy_push(1)
        s.run(parts[2])
        got = (s.raster(), s.get_vars(), s.counters())
    _same(got, want)


def test_network_without_pull_layout_rejects_a_threshold(ctx):
    import dataclasses
    from szsnn_amd import configs
    from szsnn_amd._lib import SzsnnError
    from szsnn_amd.simulation import Simulation
    s = Simulation(ctx, dataclasses.replace(configs.izhikevich2003(), pull_layout=-1))
    assert s.heavy_push() == 0
    s.set_heavy_push(0)
    with pytest.raises(SzsnnError):
        s.set_heavy_push(100)


def test_forced_burst_through_step_matches_oracle(ctx):
    """Bursts of 250 contiguous ids forced through single steps (every third
    step, ids 300..549: 25k synapse entries, posted as heavy pushes at the
    default threshold) against the FP32 oracle given the same forced ids."""
    from oracle.pyoracle import Oracle
    from szsnn_amd import configs
    from szsnn_amd.simulation import Simulation
    net, steps = configs.izhikevich2003(), 120
    forced = [(t, i) for t in range(0, steps, 3) for i in range(300, 550)]
    by_t = {}
    for t, i in forced:
        by_t.setdefault(t, []).append(i)
    print("case: forced bursts through step(), default heavy threshold", flush=True)
    s = Simulation(ctx, net)
    assert s.heavy_push() > 0 and 250 * 100 >= s.heavy_push()
    s.set_gather_policy(2.0)
    s.set_recording(0, steps)
    for t in range(steps):
        s.step(by_t.get(t))
    r, v, c = s.raster(), s.get_vars(), s.counters()
    o = Oracle(net, 32)
    ro, _ = o.run(steps, forced=forced)
    bad = np.flatnonzero((r != ro).any(axis=1))
    assert bad.size == 0, f"raster differs first at step {bad[0]}"
    np.testing.assert_array_equal(v, o.vars().astype(np.float32))
    assert c["events"] == o.events > steps // 3 * 250 * 100
