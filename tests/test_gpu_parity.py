"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Bars (integer/bitmap results bit-exact; floating point as stated):
  * FP32 oracle  : spike rasters, final state variables, synaptic-event and
                   drive-event counts identical, over whole runs.
  * FP64 oracle  : rasters identical over a stated window; voltages within a
                   stated tolerance over that window.  Windows were measured
                   as the first FP32/FP64 divergence of each network on the
                   CPU (Izhikevich 2003 net: step 296; IF nets: none in 2000).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

IZH_WINDOW = 250          # steps with identical rasters vs FP64 (measured divergence at 296)
IZH_V_TOL_MAX = 1.0       # mV, worst sample (spike upstroke amplifies 0.04 v^2)
IZH_V_TOL_P999 = 1e-2     # mV, 99.9th percentile
IF_WINDOW = 2000
IF_V_TOL = 5e-3           # mV


@pytest.fixture(scope="module")
def ctx():
    from szsnn_amd.simulation import Context
    c = Context(0)
    print("device:", c.name, flush=True)
    return c


def _gpu_run(ctx, net, steps, forced=None):
    from szsnn_amd.simulation import Simulation
    sim = Simulation(ctx, net)
    sim.set_recording(0, steps)
    if forced:
        by_t = {}
        for t, i in forced:
            by_t.setdefault(t, []).append(i)
        for t in range(steps):
            sim.step(by_t.get(t))
    else:
        sim.run(steps)
    return sim, sim.raster(), sim.get_vars(), sim.counters()


def _oracle_run(net, steps, precision, forced=None, record_v=False):
    from oracle.pyoracle import Oracle
    o = Oracle(net, precision)
    r, v = o.run(steps, forced=forced, record_v=record_v)
    return o, r, v


def _assert_fp32_exact(ctx, net, steps, forced=None):
    sim, rg, vg, cg = _gpu_run(ctx, net, steps, forced)
    o, ro, _ = _oracle_run(net, steps, 32, forced)
    bad = np.flatnonzero((rg != ro).any(axis=1))
    assert bad.size == 0, f"raster differs first at step {bad[0]}"
    vo = o.vars().astype(np.float32)
    np.testing.assert_array_equal(vg, vo)
    assert cg["events"] == o.events
    assert cg["drive_events"] == o.drive_events
    assert cg["cycle"] == steps
    return sim, rg


def test_izhikevich2003_fp32_bit_exact(ctx):
    from szsnn_amd.configs import izhikevich2003
    net = izhikevich2003()
    print("case: izhikevich2003 1000 steps", flush=True)
    _, r = _assert_fp32_exact(ctx, net, 1000)
    assert r.any()


def test_if_curr_exp_10k_fp32_bit_exact(ctx):
    from szsnn_amd.configs import if_curr_exp_10k
    net = if_curr_exp_10k()
    print("case: if_curr_exp_10k 2000 steps", flush=True)
    _assert_fp32_exact(ctx, net, 2000)


def test_if_cond_exp_scaled_fp32_bit_exact(ctx):
    from szsnn_amd.configs import if_cond_exp_100k
    net = if_cond_exp_100k(n_exc=8000, n_inh=2000, k_exc=400, k_inh=100)
    print("case: if_cond_exp 10k/500 2000 steps", flush=True)
    _assert_fp32_exact(ctx, net, 2000)


def test_izhikevich_fp64_window(ctx):
    from szsnn_amd.configs import izhikevich2003
    net = izhikevich2003()
    _, rg, _, _ = _gpu_run(ctx, net, IZH_WINDOW)
    _, ro, vo = _oracle_run(net, IZH_WINDOW, 64, record_v=True)
    np.testing.assert_array_equal(rg, ro)
    # voltages per step: rerun the GPU step by step to read v
    from szsnn_amd.simulation import Simulation
    sim = Simulation(ctx, net)
    dv = []
    for t in range(IZH_WINDOW):
        sim.step()
        dv.append(np.abs(sim.get_vars()[0].astype(np.float64) - vo[t]))
    dv = np.concatenate(dv)
    assert dv.max() <= IZH_V_TOL_MAX
    assert np.percentile(dv, 99.9) <= IZH_V_TOL_P999


@pytest.mark.parametrize("which", ["curr", "cond"])
def test_if_fp64_window(ctx, which):
    from szsnn_amd import configs
    if which == "curr":
        net = configs.if_curr_exp_10k()
    else:
        net = configs.if_cond_exp_100k(n_exc=1600, n_inh=400, k_exc=160, k_inh=40)
    _, rg, vg, _ = _gpu_run(ctx, net, IF_WINDOW)
    o, ro, _ = _oracle_run(net, IF_WINDOW, 64)
    np.testing.assert_array_equal(rg, ro)
    assert np.abs(vg[0].astype(np.float64) - o.vars()[0]).max() <= IF_V_TOL


def test_forced_input_and_const_rate_sources(ctx):
    """szsnn_step with forced ids (neurons and sources, also in refractory
    periods) and Const_rate sources, against the FP32 oracle."""
    from szsnn_amd.network import Network, if_curr_exp_params, if_vars, quantize_weights
    rng = np.random.default_rng(5)
    n, m, D, S = 300, 20, 9, 6000
    src = rng.integers(0, n + m, S).astype(np.uint32)
    tgt = rng.integers(0, n, S).astype(np.uint32)
    w = quantize_weights(rng.uniform(-0.6, 1.0, S), 26)
    dl = rng.integers(0, D, S).astype(np.uint16)
    net = Network(model=0, n_neurons=n, params=if_curr_exp_params(n, tau_refrac=3.0, v_thresh=-55.0),
                  vars=if_vars(n, rng.uniform(-65, -55, n).astype(np.float32)), syn_src=src, syn_tgt=tgt,
                  syn_w=w, syn_delay=dl, max_delay=D, frac_bits=26, seed=9, n_sources=m, source_model=1,
                  src_rate=rng.integers(2, 9, m).astype(np.float32), src_last=np.zeros(m, np.float32))
    forced = [(int(t), int(i)) for t, i in zip(rng.integers(0, 200, 400), rng.integers(0, n + m, 400))]
    _assert_fp32_exact(ctx, net, 200, forced=forced)


def test_delay_convention_single_synapse(ctx):
    """A forced spike at step c through one synapse of API delay k first
    changes the target in the update of step c+k+1."""
    from szsnn_amd.network import Network, if_curr_exp_params, if_vars
    from szsnn_amd.simulation import Simulation
    for k in (0, 1, 5, 63):
        net = Network(model=0, n_neurons=2, params=if_curr_exp_params(2, v_thresh=1e9), vars=if_vars(2),
                      syn_src=np.array([0], np.uint32), syn_tgt=np.array([1], np.uint32),
                      syn_w=np.array([0.5], np.float32), syn_delay=np.array([k], np.uint16),
                      max_delay=64, frac_bits=30)
        sim = Simulation(ctx, net)
        c = 3
        ie = []
        for t in range(c + k + 4):
            sim.step([0] if t == c else None)
            ie.append(sim.get_vars()[1, 1])
        first = int(np.flatnonzero(np.array(ie) != 0)[0])
        assert first == c + k + 1, (k, first)
        assert ie[first] == np.float32(0.5)


def test_graph_run_equals_step_loop(ctx):
    """The graph-captured loop and single steps give the same raster."""
    from szsnn_amd.configs import izhikevich2003
    from szsnn_amd.simulation import Simulation
    net = izhikevich2003(n_exc=400, n_inh=100)
    a = Simulation(ctx, net)
    a.set_recording(0, 137)
    a.run(137)
    b = Simulation(ctx, net)
    b.set_recording(0, 137)
    for _ in range(137):
        b.step()
    np.testing.assert_array_equal(a.raster(), b.raster())
    np.testing.assert_array_equal(a.get_vars(), b.get_vars())


def test_step_and_fired_ids(ctx):
    from szsnn_amd.configs import izhikevich2003
    from szsnn_amd.simulation import Simulation, bitmap_to_ids
    net = izhikevich2003(n_exc=400, n_inh=100)
    a = Simulation(ctx, net)
    a.set_recording(0, 60)
    a.run(60)
    ref = a.raster()
    b = Simulation(ctx, net)
    for t in range(60):
        ids = b.stepAndFired()
        exp = bitmap_to_ids(ref[t])
        np.testing.assert_array_equal(ids, exp[exp < net.n_neurons])
        assert np.all(np.diff(ids.astype(np.int64)) > 0)


def test_errors_are_reported(ctx):
    from szsnn_amd import SzsnnError
    from szsnn_amd.configs import izhikevich2003
    from szsnn_amd.simulation import Simulation
    net = izhikevich2003(n_exc=80, n_inh=20, fanout=10)
    sim = Simulation(ctx, net)
    with pytest.raises(SzsnnError, match="not in the network"):
        sim.step([net.n_neurons + 5])
    bad = izhikevich2003(n_exc=80, n_inh=20, fanout=10)
    bad.syn_delay[0] = bad.max_delay
    with pytest.raises(SzsnnError, match="max_delay"):
        Simulation(ctx, bad)
    with pytest.raises(SzsnnError, match="update_block"):
        sim.set_tiling(update_block=192)
    with pytest.raises(SzsnnError, match="persistent block must be 256 or 512"):
        sim.set_mode(1, 1024, 0)
    with pytest.raises(SzsnnError, match="update_ids_per_lane"):
        sim.set_tiling_ex(update_ids_per_lane=3)
    with pytest.raises(SzsnnError, match="gather_batch"):
        sim.set_tiling_ex(gather_batch=3)
    # the reference's queues could overflow (reported, reference D4); delivery
    # through the ring of accumulator slots has no queue: a tiny
    # queue_capacity is ignored and a full volley still delivers every event
    runs = []
    for cap in (0, 1):
        tiny = izhikevich2003(n_exc=80, n_inh=20, fanout=10)
        tiny.queue_capacity = cap
        sim = Simulation(ctx, tiny)
        sim.set_gather_policy(2.0)   # push every step
        sim.set_volley(100)
        sim.run(3)
        runs.append(sim.counters())
    # every id fires at steps 0, 1, 2: a synapse of delay k is delivered at the
    # steps c + k + 1 <= 2 that are inside the run, i.e. max(0, 2 - k) times
    w_q = np.round(tiny.syn_w.astype(np.float64) * 2.0 ** tiny.frac_bits)
    k = tiny.syn_delay.astype(np.int64)[w_q != 0]
    expected = int(np.maximum(0, 2 - k).sum())
    assert runs[0]["events"] == runs[1]["events"] == expected
    assert runs[1]["overflow"] == 0


def test_autotune_restores_state(ctx):
    from szsnn_amd.configs import izhikevich2003
    from szsnn_amd.simulation import Simulation
    net = izhikevich2003(n_exc=400, n_inh=100)
    a = Simulation(ctx, net)
    a.run(30)
    tiling = a.autotune(steps=50)
    assert tiling["update_block"] % 64 == 0
    a.set_recording(30, 100)
    a.run(100)
    b = Simulation(ctx, net)
    b.set_recording(30, 100)
    b.run(130)
    np.testing.assert_array_equal(a.raster(), b.raster())
    assert a.counters()["cycle"] == 130


def test_full_size_if_cond_exp_event_checksum(ctx):
    """Full 100k-neuron, 1e8-synapse network: the device's event counter equals
    the events implied by its own raster and the synapse table (each spike of
    id s at step c delivers row (s, k) at step c+k+1 if that is inside the run)."""
    from szsnn_amd.configs import if_cond_exp_100k
    from szsnn_amd.simulation import Simulation, raster_to_bool
    net = if_cond_exp_100k()
    print("case: if_cond_exp_100k full size, 300 steps", flush=True)
    sim = Simulation(ctx, net)
    assert sim.info.n_syn_stored == 100_000_000
    T = 300
    sim.set_recording(0, T)
    sim.set_volley(0)
    sim.run(T)
    spikes = raster_to_bool(sim.raster(), net.n_neurons)
    D = net.max_delay
    rowlen = np.bincount(net.syn_src.astype(np.int64) * D + net.syn_delay, minlength=net.n_neurons * D)
    rowlen = rowlen.reshape(net.n_neurons, D)
    exp = 0
    for c in range(T):
        ids = np.flatnonzero(spikes[c])
        if ids.size == 0:
            continue
        kmax = T - c - 1  # deliveries at step c+k+1 <= T-1
        if kmax <= 0:
            continue
        exp += int(rowlen[ids, :min(D, kmax)].sum())
    cnt = sim.counters()
    assert cnt["events"] == exp
    assert cnt["spikes"] == int(spikes.sum())
    assert cnt["overflow"] == 0


@pytest.mark.parametrize("block,grid", [(256, 0), (512, 16), (512, 4)])
def test_persistent_mode_equals_graph_mode(ctx, block, grid):
    """The cooperative persistent step loop gives the same rasters and state
    as the graph-captured per-kernel loop."""
    from szsnn_amd._lib import MODE_PERSISTENT
    from szsnn_amd.configs import if_curr_exp_10k, izhikevich2003
    from szsnn_amd.simulation import Simulation
    for net, steps in ((izhikevich2003(n_exc=1600, n_inh=400), 300),
                       (if_curr_exp_10k(n_exc=2000, n_inh=500, n_poisson=300, poisson_hz=100.0,
                                        w_poisson=1.5), 500)):
        a = Simulation(ctx, net)
        a.set_recording(0, steps)
        a.run(steps)
        b = Simulation(ctx, net)
        b.set_mode(MODE_PERSISTENT, block, grid)
        b.set_recording(0, steps)
        b.run(steps // 3)
        b.run(steps - steps // 3)
        np.testing.assert_array_equal(a.raster(), b.raster())
        np.testing.assert_array_equal(a.get_vars(), b.get_vars())
        ca, cb = a.counters(), b.counters()
        assert ca["events"] == cb["events"] and cb["cycle"] == steps


def _run_policy(ctx, net, steps, fraction, volley=0):
    from szsnn_amd.simulation import Simulation
    s = Simulation(ctx, net)
    assert s.gather_policy()["has_pull"]
    s.set_gather_policy(fraction)
    s.set_volley(volley)
    s.set_recording(0, steps)
    s.run(steps)
    return s.raster(), s.get_vars(), s.counters()


def _mixed_fraction(net, raster, volley):
    """A policy fraction under which about half of the steps are deferred:
    step t is deferred when its predicted spikes, spikes(t-1) less the
    volley forced in step t-1 plus the volley of step t, reach ceil(fraction
    x ids with synapses) (the volley runs from step 0).  Returns (fraction,
    deferred steps)."""
    from szsnn_amd.simulation import raster_to_bool
    counts = raster_to_bool(raster, net.n_neurons + net.n_sources).sum(axis=1).astype(np.int64)
    pred = np.concatenate([[volley], counts[:-1]])
    active = np.unique(net.syn_src[net.syn_w != 0]).size
    thr = int(np.median(pred[1:])) + 1
    frac = (thr - 0.5) / active
    return frac, int((pred >= thr).sum())


@pytest.mark.parametrize("which", ["izh", "curr", "cond_volley", "cond_volley_6", "cond_volley_8"])
def test_pull_gather_equals_push_gather(ctx, which):
    """Dense (pull) and sparse (push) gathers give identical sums, so rasters,
    state and event counts agree bit for bit; pull-only also matches the oracle.
    The cond_volley cases differ in size, so the upload picks 4, 6 and 8
    slices of 64 targets per pull block (on an MI355X: 256 CUs, 7 gather
    blocks each)."""
    from szsnn_amd import configs
    from szsnn_amd.simulation import Simulation
    volley, slices = 0, None
    if which == "izh":
        net, steps = configs.izhikevich2003(), 300
    elif which == "curr":
        net, steps = configs.if_curr_exp_10k(n_exc=2000, n_inh=500, n_poisson=300, poisson_hz=100.0,
                                             w_poisson=1.5), 400
    elif which == "cond_volley":
        net, steps, volley = configs.if_cond_exp_100k(n_exc=4000, n_inh=1000, k_exc=200, k_inh=50), 60, 1000
        slices = 4
    elif which == "cond_volley_6":
        net, steps, volley = configs.if_cond_exp_100k(n_exc=8800, n_inh=2200, k_exc=200, k_inh=50), 60, 2200
        slices = 6
    else:
        net, steps, volley = configs.if_cond_exp_100k(n_exc=12800, n_inh=3200, k_exc=200, k_inh=50), 60, 3200
        slices = 8
    s = Simulation(ctx, net)
    got = int(s.info.pull_slices)
    s.close()
    print(f"case: pull vs push {which}, {got} slices per pull block", flush=True)
    if slices is not None and "MI355" in ctx.name:
        assert got == slices
    rp, vp, cp = _run_policy(ctx, net, steps, 2.0, volley)   # push only
    rq, vq, cq = _run_policy(ctx, net, steps, 0.0, volley)   # pull only
    frac, n_def = _mixed_fraction(net, rp, volley)
    assert 0 < n_def < steps, (frac, n_def)
    ra, va, ca = _run_policy(ctx, net, steps, frac, volley)  # some steps deferred
    for r, v, c in ((rq, vq, cq), (ra, va, ca)):
        np.testing.assert_array_equal(r, rp)
        np.testing.assert_array_equal(v, vp)
        assert c["events"] == cp["events"] and c["spikes"] == cp["spikes"]
    if volley == 0:
        o, ro, _ = _oracle_run(net, steps, 32)
        np.testing.assert_array_equal(rq, ro)
        assert cq["events"] == o.events


def test_persistent_mode_pull_volley_equals_graph(ctx):
    """The persistent kernel's pull path (dense gather inside the cooperative
    loop) against the graph mode, under a volley that makes every step pull."""
    from szsnn_amd._lib import MODE_PERSISTENT
    from szsnn_amd.configs import if_cond_exp_100k
    from szsnn_amd.simulation import Simulation
    net = if_cond_exp_100k(n_exc=4000, n_inh=1000, k_exc=200, k_inh=50)
    steps = 80
    runs = []
    for mode in (0, MODE_PERSISTENT):
        s = Simulation(ctx, net)
        if mode:
            s.set_mode(MODE_PERSISTENT, 256, 0)
        s.set_gather_policy(0.0)   # pull every step
        s.set_volley(2500)
        s.set_recording(0, steps)
        s.run(steps)
        runs.append((s.raster(), s.get_vars(), s.counters()))
    np.testing.assert_array_equal(runs[0][0], runs[1][0])
    np.testing.assert_array_equal(runs[0][1], runs[1][1])
    assert runs[0][2]["events"] == runs[1][2]["events"] > 0


@pytest.mark.parametrize("which", ["count", "count_skips", "table", "long_lanes"])
def test_narrow_and_wide_pull_entries_agree(ctx, which):
    """The same network pulled through narrow entries (2-byte count format:
    few weights, sources as running deltas, gaps wider than a delta taken by
    skip entries; or the 4-byte table format, up to 255 weights looked up)
    and through 8-byte entries: rasters, state and event counts identical to
    pushing every step; the pulled table network also matches the FP32 oracle.
    long_lanes: 1850 synapses per target, all of one delay, so every slice
    lane's column is longer than its 16-bit lane word can count (the lane
    then reads all the slice's vectors, its padding included)."""
    import dataclasses
    from szsnn_amd import configs
    from szsnn_amd.simulation import Simulation
    if which == "count":
        net, steps, volley = configs.if_cond_exp_100k(n_exc=4000, n_inh=1000, k_exc=200, k_inh=50), 70, 1500
    elif which == "count_skips":
        # 50000 sources, about 8 per (target, delay): gaps beyond 2^14 - 1 are common
        net, steps, volley = configs.if_cond_exp_100k(n_exc=40000, n_inh=10000, k_exc=200, k_inh=50), 40, 12000
    elif which == "long_lanes":
        net, steps, volley = configs.if_cond_exp_100k(n_exc=1600, n_inh=400, k_exc=1500, k_inh=350,
                                                      max_delay_steps=1), 40, 600
        per_target = np.bincount(net.syn_tgt[net.syn_w != 0], minlength=net.n_neurons)
        assert per_target.min() > 8 * 127  # 2-byte entries: more than 127 vectors per lane
    else:
        net, steps, volley = configs.izhikevich2003(), 300, 0
        w = net.syn_w.copy()
        net = dataclasses.replace(net, syn_w=(np.round(w * 25.0) / 25.0).astype(np.float32))
        assert 20 < np.unique(net.syn_w[net.syn_w != 0]).size <= 255
    print(f"case: narrow vs wide pull entries ({which})", flush=True)
    narrow_bytes = 4 if which == "table" else 2
    default_fraction = {2: 0.012, 4: 0.02, 8: 0.04}   # where a pushed step costs as much as a pull
    for layout, width in ((0, narrow_bytes), (1, 8)):
        s = Simulation(ctx, dataclasses.replace(net, pull_layout=layout))
        assert s.info.pull_entry_bytes == width
        assert s.gather_policy()["dense_fraction"] == pytest.approx(default_fraction[width])
        del s
    rp, vp, cp = _run_policy(ctx, net, steps, 2.0, volley)                                    # push only
    rn, vn, cn = _run_policy(ctx, net, steps, 0.0, volley)                                    # pull, narrow
    r8, v8, c8 = _run_policy(ctx, dataclasses.replace(net, pull_layout=1), steps, 0.0, volley)  # pull, 8 B
    for r, v, c in ((rn, vn, cn), (r8, v8, c8)):
        np.testing.assert_array_equal(r, rp)
        np.testing.assert_array_equal(v, vp)
        assert c["events"] == cp["events"] > 0
    if volley == 0:
        o, ro, _ = _oracle_run(net, steps, 32)
        np.testing.assert_array_equal(rn, ro)
        assert cn["events"] == o.events


@pytest.mark.parametrize("tiling", [
    {"gather_stage": 0, "gather_batch": 2, "gather_block": 256},
    {"gather_stage": 0, "gather_batch": 4, "gather_block": 512},
    {"gather_stage": 0, "gather_batch": 2, "gather_block": 1024},
    {"gather_stage": 1, "gather_batch": 4, "gather_block": 256},
    {"gather_stage": 1, "gather_batch": 2, "gather_block": 1024},
])
@pytest.mark.parametrize("which", ["izh", "cond_volley"])
def test_pull_tilings_equal_push(ctx, which, tiling):
    """The pull's launch shapes (several delays staged per pass, batches of
    4 vectors, more waves than slices) give the push's rasters, state and
    event counts, also when only some steps are deferred."""
    from szsnn_amd import configs
    from szsnn_amd.simulation import Simulation
    volley = 0
    if which == "izh":
        net, steps = configs.izhikevich2003(), 300
    else:
        net, steps, volley = configs.if_cond_exp_100k(n_exc=4000, n_inh=1000, k_exc=200, k_inh=50), 60, 1000
    print(f"case: pull tiling {tiling} on {which}", flush=True)
    rp, vp, cp = _run_policy(ctx, net, steps, 2.0, volley)   # push only
    frac, n_def = _mixed_fraction(net, rp, volley)
    for fraction in (0.0, frac):
        s = Simulation(ctx, net)
        s.set_tiling_ex(**tiling)
        t = s.tiling_ex()
        assert t["gather_stage"] == (t["gather_stage_max"] if tiling["gather_stage"] == 0 else 1)
        s.set_gather_policy(fraction)
        s.set_volley(volley)
        s.set_recording(0, steps)
        s.run(steps)
        np.testing.assert_array_equal(s.raster(), rp)
        np.testing.assert_array_equal(s.get_vars(), vp)
        c = s.counters()
        assert c["events"] == cp["events"] and c["rows"] == cp["rows"]
        s.close()


def test_update_ids_per_lane_equal_one(ctx):
    """Update blocks serving 2 or 4 id blocks (ids per lane) give the same
    run as one id per lane, in graph mode with pushes and with pulls."""
    from szsnn_amd.configs import if_curr_exp_10k, izhikevich2003
    from szsnn_amd.simulation import Simulation
    for net, steps in ((izhikevich2003(n_exc=1600, n_inh=400), 300),
                       (if_curr_exp_10k(n_exc=2000, n_inh=500, n_poisson=300, poisson_hz=100.0, w_poisson=1.5), 400)):
        runs = []
        for npl, block in ((1, 256), (2, 256), (4, 128), (4, 1024)):
            s = Simulation(ctx, net)
            s.set_tiling_ex(update_block=block, update_ids_per_lane=npl)
            s.set_recording(0, steps)
            s.run(steps)
            runs.append((s.raster(), s.get_vars(), s.counters()))
            s.close()
        for r in runs[1:]:
            np.testing.assert_array_equal(r[0], runs[0][0])
            np.testing.assert_array_equal(r[1], runs[0][1])
            for k in ("events", "spikes", "rows", "drive_events"):
                assert r[2][k] == runs[0][2][k], k
