"""GPU edge cases against the FP32 oracle (bit for bit): a network without
synapses, the largest delay ring (64 slots) with ids that do not fill whole
wavefronts, and pulls at the largest delay."""
import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from szsnn_amd.simulation import Context
    return Context(0)


def _ragged_net(seed=7, n=1001, m=37, max_delay=64, fanout=40):
    """IF_curr_exp, N = 1001 neurons and 37 Poisson sources (ids not a multiple
    of 64), delays over all 64 slots, three distinct weights."""
    from szsnn_amd._lib import MODEL_IF_CURR_EXP, SOURCE_POISSON
    from szsnn_amd.network import (Network, choose_frac_bits, if_curr_exp_params, if_vars, poisson_thresholds,
                                   quantize_weights)
    rng = np.random.default_rng(seed)
    T = n + m
    src = np.repeat(np.arange(T, dtype=np.uint32), fanout)
    tgt = rng.integers(0, n, src.size).astype(np.uint32)
    w = rng.choice(np.array([0.6, -1.2, 2.5], np.float32), src.size)
    dl = rng.integers(0, max_delay, src.size).astype(np.uint16)
    dl[:T] = max_delay - 1  # every id has a synapse of the largest delay
    fb = choose_frac_bits(float(np.abs(w).max()))
    return Network(model=MODEL_IF_CURR_EXP, n_neurons=n, params=if_curr_exp_params(n, dt=1.0, v_rest=-65.0),
                   vars=if_vars(n), syn_src=src, syn_tgt=tgt, syn_w=quantize_weights(w, fb), syn_delay=dl,
                   max_delay=max_delay, frac_bits=fb, seed=seed, n_sources=m, source_model=SOURCE_POISSON,
                   src_thresh=poisson_thresholds(np.full(m, 0.05))).validate()


def _run(ctx, net, steps, fraction=None, volley=0, heavy=None):
    from szsnn_amd.simulation import Simulation
    s = Simulation(ctx, net)
    if fraction is not None:
        s.set_gather_policy(fraction)
    if heavy is not None:
        s.set_heavy_push(heavy)
    if volley:
        s.set_volley(volley)
    s.set_recording(0, steps)
    s.run(steps)
    return s.raster(), s.get_vars(), s.counters(), s


def test_network_without_synapses(ctx):
    """No synapses: nothing to push or pull; the Gaussian-driven neurons still
    match the oracle, and no synaptic event is counted."""
    from oracle.pyoracle import Oracle
    from szsnn_amd.configs import izhikevich2003
    net = izhikevich2003(n_exc=300, n_inh=77)
    empty = np.zeros(0, np.uint32)
    net = dataclasses.replace(net, syn_src=empty, syn_tgt=empty.copy(), syn_w=np.zeros(0, np.float32),
                              syn_delay=np.zeros(0, np.uint16))
    print("case: network without synapses", flush=True)
    r, v, c, s = _run(ctx, net, 200)
    assert s.info.pull_entry_bytes == 0
    o = Oracle(net, 32)
    ro, _ = o.run(200)
    np.testing.assert_array_equal(r, ro)
    np.testing.assert_array_equal(v, o.vars().astype(np.float32))
    assert c["events"] == 0 and o.events == 0 and c["spikes"] > 0


@pytest.mark.parametrize("fraction,heavy", [(2.0, None), (2.0, 1), (2.0, 0), (0.0, None)])
def test_largest_delay_ring_ragged_ids(ctx, fraction, heavy):
    """64 delay slots, 1001 neurons + 37 sources, a volley of 600 forced ids:
    pushing every step (by the update blocks, as heavy pushes at the default
    threshold, or as heavy pushes from every block) and pulling every step
    (2-byte entries) all equal the oracle bit for bit, including synapses of
    delay 63, whose delivery slots wrap around the ring."""
    from oracle.pyoracle import Oracle
    net = _ragged_net()
    steps, volley = 140, 600
    print(f"case: largest delay ring, policy {fraction}, heavy threshold {heavy}", flush=True)
    r, v, c, s = _run(ctx, net, steps, fraction=fraction, volley=volley, heavy=heavy)
    assert s.info.pull_entry_bytes == 2 and s.info.max_delay == 64
    forced = [(t, i) for t in range(steps) for i in range(volley)]
    o = Oracle(net, 32)
    ro, _ = o.run(steps, forced=forced)
    bad = np.flatnonzero((r != ro).any(axis=1))
    assert bad.size == 0, f"raster differs first at step {bad[0] if bad.size else -1}"
    np.testing.assert_array_equal(v, o.vars().astype(np.float32))
    assert c["events"] == o.events > 0


def test_persistent_then_single_steps(ctx):
    """A persistent launch followed by single steps with forced ids (state,
    spike counts and deferral facts hand over between the two kinds of
    launch) equals the same steps taken one at a time."""
    from szsnn_amd._lib import MODE_PERSISTENT
    from szsnn_amd.simulation import Simulation
    net = _ragged_net(seed=11, max_delay=16)
    runs = []
    for persistent in (False, True):
        s = Simulation(ctx, net)
        s.set_gather_policy(0.02)
        if persistent:
            s.set_mode(MODE_PERSISTENT, 256, 0)
        s.set_recording(0, 160)
        s.run(100)
        for t in range(60):
            s.step([t % 1001, (7 * t) % 1038] if t % 3 == 0 else None)
        runs.append((s.raster(), s.get_vars(), s.counters()))
    np.testing.assert_array_equal(runs[0][0], runs[1][0])
    np.testing.assert_array_equal(runs[0][1], runs[1][1])
    assert runs[0][2]["events"] == runs[1][2]["events"]
    assert runs[0][2]["spikes"] == runs[1][2]["spikes"]


@pytest.mark.parametrize("last", ["heavy", "policy"])
def test_gather_left_out_only_after_the_last_due_pull(ctx, last):
    """With no heavy pushes and a policy that never defers, steps leave the
    gather launch out, but only once no deferred step can still be due.
    20 steps pulled every step (each deferred; delays over all 64 slots),
    then 20 steps with the last switch still pending: heavy pushes from
    every block (then heavy pushes off) or still pulling (then a policy
    that never defers).  The following graph and single steps, the first R
    of them still pulling the deferred steps' late delays, equal the FP32
    oracle bit for bit."""
    from oracle.pyoracle import Oracle
    from szsnn_amd.simulation import Simulation
    net = _ragged_net(seed=13)
    s = Simulation(ctx, net)
    s.set_volley(200)
    steps = 40 + 150 + 7
    s.set_recording(0, steps)
    print(f"case: gather left out after the last due pull (last switch: {last})", flush=True)
    s.set_heavy_push(0)
    s.set_gather_policy(0.0)
    s.run(20)
    if last == "heavy":
        s.set_gather_policy(2.0)
        s.set_heavy_push(1)
        s.run(20)
        s.set_heavy_push(0)
    else:
        s.run(20)
        s.set_gather_policy(2.0)
    s.run(150)       # three 50-step graphs: gathers in the first two, none in the last
    for _ in range(7):
        s.step()
    r, v, c = s.raster(), s.get_vars(), s.counters()
    s.close()
    o = Oracle(net, 32)
    ft = np.repeat(np.arange(steps, dtype=np.uint64), 200)
    fi = np.tile(np.arange(200, dtype=np.uint32), steps)
    ro, _ = o.run(steps, forced=(ft, fi))
    bad = np.flatnonzero((r != ro).any(axis=1))
    assert bad.size == 0, f"raster differs first at step {bad[0]}"
    np.testing.assert_array_equal(v, o.vars().astype(np.float32))
    assert c["events"] == o.events and c["events"] > 0
