"""GPU parity at the configurations' real sizes, through the C ABI.

  * IF_cond_exp 100k network (1e8 synapses, 32 delay slots): its whole
    configured run, 10000 steps, bit-exact against the FP32 oracle
    (rasters, final state, synaptic and drive event counts);
  * IF_curr_exp 10k network: 10000 steps against the FP32 oracle, as
    configured and with a stronger Poisson drive that makes it fire;
  * the batched Izhikevich configuration (256 instances of 500..20000
    neurons, 100 synapses each, noise amplitudes U[2,8]): every instance
    equals its plain run, and the smallest, the largest and two more equal
    the FP32 oracle, over 1000 steps;
  * the volley workload the roofline is quoted on (every neuron of the 100k
    network forced each step, 2-byte pull entries, gather grid 4096): pulled
    and pushed steps give the same rasters, state and event counts, and both
    equal the FP32 oracle;
  * persistent mode with grids that do not give every id a lane (the loop
    over id blocks inside each block) equal to graph mode.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from szsnn_amd.simulation import Context
    c = Context(0)
    print("device:", c.name, flush=True)
    return c


@pytest.fixture(scope="module")
def net100k():
    from szsnn_amd.configs import if_cond_exp_100k
    return if_cond_exp_100k()


def _gpu(ctx, net, steps, volley=0, volley_steps=None, policy=None, tiling=None, mode=None):
    """Run on the device; returns (raster, vars, counters).  A volley forces
    ids [0, volley) in the first volley_steps steps."""
    from szsnn_amd.simulation import Simulation
    s = Simulation(ctx, net)
    if policy is not None:
        s.set_gather_policy(policy)
    if tiling:
        s.set_tiling(**tiling)
    if mode:
        s.set_mode(*mode)
    s.set_recording(0, steps)
    if volley:
        n1 = steps if volley_steps is None else volley_steps
        s.set_volley(volley)
        s.run(n1)
        s.set_volley(0)
        if steps > n1:
            s.run(steps - n1)
    else:
        s.run(steps)
    out = (s.raster(), s.get_vars(), s.counters())
    s.close()
    return out


def _oracle(net, steps, forced=None):
    """The FP32 oracle on the host's cores (its parallel step equals the
    serial one bit for bit)."""
    from oracle.pyoracle import Oracle, host_threads
    o = Oracle(net, 32)
    r, _ = o.run(steps, forced=forced, threads=host_threads())
    return r, o.vars().astype(np.float32), o.events, o.drive_events


def _assert_same(gpu, ref, what):
    rg, vg, cg = gpu
    ro, vo, eo, do = ref
    bad = np.flatnonzero((rg != ro).any(axis=1))
    assert bad.size == 0, f"{what}: raster differs first at step {bad[0]}"
    np.testing.assert_array_equal(vg, vo, err_msg=what)
    assert cg["events"] == eo, (what, cg["events"], eo)
    assert cg["drive_events"] == do, (what, cg["drive_events"], do)
    assert cg["overflow"] == 0


def test_if_cond_exp_100k_full_run_fp32_bit_exact(ctx, net100k):
    from szsnn_amd.simulation import raster_to_bool
    steps = 10000
    print("case: if_cond_exp_100k full size, 10000 steps, vs FP32 oracle", flush=True)
    gpu = _gpu(ctx, net100k, steps)
    ref = _oracle(net100k, steps)
    _assert_same(gpu, ref, "if_cond_exp_100k")
    spikes = raster_to_bool(gpu[0], net100k.n_neurons).sum()
    assert spikes > 10000 and gpu[2]["events"] > 1e7, (spikes, gpu[2]["events"])
    assert gpu[2]["spikes"] == spikes and gpu[2]["cycle"] == steps


@pytest.mark.parametrize("drive", ["as_configured", "driven"])
def test_if_curr_exp_10k_10000_steps_fp32_bit_exact(ctx, drive):
    """As configured, the network's neurons stay below threshold (five 10 Hz
    inputs of 0.1 nA per neuron: about 0.5 mV of mean depolarisation), so
    only its Poisson sources spike; the driven variant (100 Hz sources of
    0.7 nA) fires at about 57 Hz."""
    from szsnn_amd.configs import if_curr_exp_10k
    from szsnn_amd.simulation import raster_to_bool
    net = if_curr_exp_10k() if drive == "as_configured" else if_curr_exp_10k(poisson_hz=100.0, w_poisson=0.7)
    print(f"case: if_curr_exp_10k ({drive}) 10000 steps", flush=True)
    gpu = _gpu(ctx, net, 10000)
    _assert_same(gpu, _oracle(net, 10000), f"if_curr_exp_10k {drive}")
    b = raster_to_bool(gpu[0], net.n_neurons + net.n_sources)
    assert b[:, net.n_neurons:].sum() > 5000    # the Poisson sources
    if drive == "driven":
        assert b[:, :net.n_neurons].sum() > 4e5   # about 57 Hz x 10000 neurons x 1 s


def test_batched_izhikevich_256_instances(ctx):
    from oracle.pyoracle import Oracle, host_threads
    from szsnn_amd.batch import BatchSimulation, batched_izhikevich
    from szsnn_amd.simulation import Simulation, raster_to_bool
    nets = batched_izhikevich()
    sizes = np.array([n.n_neurons for n in nets])
    assert len(nets) == 256 and sizes.min() >= 500 and sizes.max() <= 20000
    steps = 1000
    print(f"case: batched izhikevich, 256 instances, {sizes.sum()} neurons, {steps} steps", flush=True)
    bs = BatchSimulation(ctx, nets)
    bs.set_recording(0, steps)
    bs.run(steps)
    r = bs.raster()
    v = bs.get_vars()
    cb = bs.counters()
    bs.close()
    events = 0
    for b, net in enumerate(nets):
        s = Simulation(ctx, net)
        s.set_recording(0, steps)
        s.run(steps)
        pr = raster_to_bool(s.raster(), net.n_neurons)
        pv = s.get_vars()
        events += s.counters()["events"]
        s.close()
        ir = bs.instance_raster(b, r)
        assert np.array_equal(ir, pr), f"instance {b} ({net.n_neurons} neurons): raster differs from its plain run"
        np.testing.assert_array_equal(bs.instance_vars(b, v), pv, err_msg=f"instance {b}")
    assert cb["events"] == events
    picks = sorted({int(sizes.argmin()), int(sizes.argmax()), 0, int(np.argsort(sizes)[len(sizes) // 2])})
    assert len(picks) >= 4
    for b in picks:
        o = Oracle(nets[b], 32)
        ro, _ = o.run(steps, threads=host_threads())
        assert np.array_equal(bs.instance_raster(b, r), raster_to_bool(ro, nets[b].n_neurons)), f"instance {b}"
        np.testing.assert_array_equal(bs.instance_vars(b, v), o.vars().astype(np.float32), err_msg=f"instance {b}")
        assert bs.instance_raster(b, r).any()


def test_full_size_volley_pull_equals_push_and_oracle(ctx, net100k):
    """Every neuron forced in the first D + 2 steps (so every delay slot is
    filled and pulled), then D + 2 free steps that drain the ring: the
    2-byte pull at gather grid 4096 against pushing every step."""
    D = net100k.max_delay
    n1, steps = D + 2, 2 * (D + 2)
    print(f"case: full-size volley, {n1} forced + {steps - n1} free steps", flush=True)
    pull = _gpu(ctx, net100k, steps, volley=net100k.n_neurons, volley_steps=n1, policy=0.0,
                tiling={"gather_block": 256, "gather_grid": 4096})
    push = _gpu(ctx, net100k, steps, volley=net100k.n_neurons, volley_steps=n1, policy=2.0)
    np.testing.assert_array_equal(pull[0], push[0])
    np.testing.assert_array_equal(pull[1], push[1])
    for k in ("events", "rows", "spikes", "drive_events"):
        assert pull[2][k] == push[2][k], k
    # every synapse is delivered once per forced step (the last forced step's
    # delay-(D-1) synapses arrive at step 2D + 1 < 2(D + 2))
    assert pull[2]["events"] >= n1 * 100_000_000
    ft = np.repeat(np.arange(n1, dtype=np.uint64), net100k.n_neurons)
    fi = np.tile(np.arange(net100k.n_neurons, dtype=np.uint32), n1)
    _assert_same(pull, _oracle(net100k, steps, forced=(ft, fi)), "volley pull")


@pytest.mark.parametrize("block,grid", [(256, 2), (512, 1)])
def test_persistent_non_covering_grid_equals_graph(ctx, block, grid):
    """Persistent launches whose grid x block is smaller than the id count:
    each block walks its id blocks every step (state loaded and stored per
    step, no carried registers)."""
    from szsnn_amd._lib import MODE_PERSISTENT
    from szsnn_amd.configs import if_cond_exp_100k, if_curr_exp_10k, izhikevich2003
    cases = [(izhikevich2003(), 400, 0, None),
             (if_curr_exp_10k(n_exc=2000, n_inh=500, n_poisson=300, poisson_hz=100.0, w_poisson=1.5), 400, 0, None),
             (if_cond_exp_100k(n_exc=4000, n_inh=1000, k_exc=200, k_inh=50), 60, 2500, 0.0)]
    for net, steps, volley, policy in cases:
        assert grid * block < net.n_neurons + net.n_sources
        print(f"case: persistent block {block} grid {grid}, {net.n_neurons} neurons, volley {volley}", flush=True)
        a = _gpu(ctx, net, steps, volley=volley, policy=policy)
        b = _gpu(ctx, net, steps, volley=volley, policy=policy, mode=(MODE_PERSISTENT, block, grid))
        np.testing.assert_array_equal(a[0], b[0])
        np.testing.assert_array_equal(a[1], b[1])
        for k in ("events", "spikes", "rows", "drive_events", "cycle"):
            assert a[2][k] == b[2][k], k
        assert a[2]["events"] > 0
