"""Benchmark: synaptic events/s and realtime factor of the plain IF_cond_exp
100k-neuron network (1e8 synapses, dt = 0.1 ms) on one MI355X.

A bench "step" is one --chunk of simulation steps (default 1000 steps of
gather -> update -> scatter, i.e. 100 ms of network time): the network fires
in population bursts of about 100 ms separated by quiet spells of about 50 ms,
so single simulation steps are not a representative unit.  Inputs (network,
state) are resident in HBM before the timed region.  The network is first run
for --settle-steps simulation steps (default 10000: the whole 1000 ms of the
configuration from its initial state, timed and reported as `full_run`), then
--warmup untimed chunks, then --steps timed chunks; the run fails if the timed
window delivers no synaptic event.
Prints one JSON line with:
  value          synaptic events/s over the timed steps (natural dynamics)
  roofline       the pull gather kernel at the volley load (all neurons forced
                 every step: every synapse delivered each step), each launch
                 preceded by a 2 GiB read sweep that evicts L2 and the
                 Infinity Cache (cold), timed with HIP events that the
                 kernel's own dispatch records on the launch stream
                 (hipExtLaunchKernelGGL: the kernel's duration, as rocprofv3
                 measures it);
                 `traffic` = HBM bytes per launch from the rocprofv3
                 FETCH_SIZE / WRITE_SIZE passes of the same cold workload
                 recorded in profiles/, `achieved` = traffic / launch time,
                 `frac` = achieved / 8 TB/s.  The survey's algorithmic unit
                 (8 B per delivered synapse + 8 B per delivered (source,
                 delay) row) is reported beside it as `alg_*`: the pull layout
                 stores 2-byte entries, so that unit is not a byte count of
                 this kernel
  cpu_baseline   the FP32 CPU restatement over the same settle run and the
                 same timed window, on every host core this process may use
                 (and a one-thread sample)
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed chunks")
    ap.add_argument("--warmup", type=int, default=2, help="untimed chunks before the timed ones")
    ap.add_argument("--chunk", type=int, default=1000, help="simulation steps per bench step")
    ap.add_argument("--settle-steps", type=int, default=10000,
                    help="untimed-for-value steps run first (the full 1000 ms network run, reported as full_run)")
    ap.add_argument("--flush-mib", type=int, default=2048,
                    help="cold-cache gather timing: MiB read between launches to evict L2 and the Infinity Cache")
    ap.add_argument("--volley-steps", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=1000, help="one-thread CPU baseline sample (steps of the same net)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--small", action="store_true", help="scaled-down network (smoke use only)")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-vs-plain comparison")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint timing")
    ap.add_argument("--batch-instances", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=1000)
    ap.add_argument("--batch-single", action="store_true",
                    help="batched pass: every step a separate launch, no step graphs (rocprofv3 --pmc "
                         "collection did not survive graph launches on this part)")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default tiling")
    ap.add_argument("--profile-pass", action="store_true",
                    help="profiler pass: a 1 GiB device copy (calibration), the volley's delay slots filled by "
                         "one persistent launch, then exactly --volley-steps gather_kernel launches")
    ap.add_argument("--cold", action="store_true",
                    help="profile pass: a --flush-mib read sweep before every gather launch")
    ap.add_argument("--fill", choices=("persistent", "graph"), default="persistent",
                    help="profile pass: fill the delay slots with one persistent (cooperative) launch, whose "
                         "gather work is inside it, or with single graph-mode steps (more gather_kernel "
                         "dispatches; no cooperative launch, after which rocprofv3 exits with SIGSEGV)")
    ap.add_argument("--natural-pass", action="store_true",
                    help="profiler pass: natural dynamics, --warmup steps then --steps steps as ONE persistent "
                         "launch (block 512, grid 256), then --volley-steps single-step gather+update launches")
    ap.add_argument("--batched-pass", action="store_true",
                    help="profiler pass: the batched Izhikevich instances only, --warmup then --batch-steps steps")
    ap.add_argument("--adjoint-pass", action="store_true",
                    help="profiler pass: the adjoint forward + backward only (2 untimed, 5 timed batches)")
    ap.add_argument("--burst-pass", action="store_true",
                    help="timing pass: ids 0..S-1 forced every step (a concentrated burst), graph mode, pushed "
                         "with heavy pushes at the default threshold, pushed by each block itself, and pulled")
    ap.add_argument("--burst-ids", default="50,250,1000,4000", help="burst pass: the values of S")
    ap.add_argument("--burst-steps", type=int, default=200, help="burst pass: timed simulation steps per case")
    ap.add_argument("--gather-block", type=int, default=256, help="profile pass: pull gather block size")
    ap.add_argument("--gather-grid", type=int, default=4096, help="profile pass: pull gather grid")
    ap.add_argument("--traffic-file", default=os.path.join(ROOT, "profiles", "gather_pull_volley_traffic.json"),
                    help="HBM bytes per pull gather launch (warm: launches back to back) measured with "
                         "rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE)")
    ap.add_argument("--cold-traffic-file",
                    default=os.path.join(ROOT, "profiles", "gather_pull_volley_cold_traffic.json"),
                    help="the same for cold launches (a cache-evicting sweep before each)")
    return ap.parse_args()


def gather_alg_bytes(events, rows):
    """Algorithmic bytes of one gather launch: an f32 weight and a u32 target
    per delivered synapse, and a (u32 start, u32 len) row pointer per
    delivered (source, delay) row (SynapseData.hpp:51-52, types.h:38)."""
    return 8 * events + 8 * rows


def copy_bandwidth_gbs(torch, nbytes=1 << 30, reps=10):
    """Device-to-device copy rate (read + write bytes / time): the practical
    HBM ceiling next to the 8 TB/s peak."""
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    b.copy_(a)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    gbs = 2 * nbytes * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del a, b
    torch.cuda.empty_cache()
    return gbs


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class CacheFlush:
    """A device buffer several times the Infinity Cache (256 MB) and the L2s;
    reading it evicts whatever a following kernel would otherwise find there.
    The sweep reads rather than writes: a write sweep leaves the Infinity
    Cache full of dirty lines, and the next kernel then pays for writing them
    back (a 253 MB read measured 104 us after a 2 GiB write sweep against 75
    us after a read sweep and 49 us warm)."""

    def __init__(self, torch, mib, stream=None):
        self.torch = torch
        self.buf = torch.ones(mib * (1 << 20) // 4, dtype=torch.float32, device="cuda")
        self.out = torch.zeros((), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        # with the library's stream, the sweep is queued ahead of the timed
        # launch on that stream, so the host's launch latency falls inside the
        # sweep and the start event fires just before the kernel starts
        self.stream = torch.cuda.ExternalStream(stream) if stream else None

    def __call__(self):
        if self.stream is None:
            self.torch.sum(self.buf, dim=0, out=self.out)
            self.torch.cuda.synchronize()
        else:
            with self.torch.cuda.stream(self.stream):
                self.torch.sum(self.buf, dim=0, out=self.out)

    def close(self):
        del self.buf
        self.torch.cuda.empty_cache()


def time_cold(sim, flush, n):
    """n single steps, each after a cache-evicting sweep queued on the
    library's stream: summed gather and update kernel times (ms, HIP events
    recorded by the kernels' dispatches on that stream)."""
    g = u = 0.0
    for _ in range(n):
        flush()
        a, b = sim.time_kernels(1)
        g += a
        u += b
    return g, u


def read_json(path):
    if path and os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return None


def run_adjoint(args, log):
    """Differentiable feedforward network (B=64, 1000 -> 2000 IF_curr_exp -> 100
    readouts, T=200, checkpoint every 20 steps): HIP forward+backward per batch
    next to the FP64 torch autograd restatement on the host CPU."""
    import numpy as np
    import torch
    from szsnn_amd.adjoint import make_problem, snn_loss
    P = make_problem()
    dev = torch.device("cuda:0")
    t = {k: torch.tensor(P[k], dtype=torch.float32, device=dev).contiguous()
         for k in ("W1", "W2", "X", "target", "params")}

    def once():
        W1 = t["W1"].clone().requires_grad_(True)
        W2 = t["W2"].clone().requires_grad_(True)
        loss, _, bits = snn_loss(W1, W2, t["X"], t["target"], t["params"], P["alpha"], K=20)
        loss.backward()
        return W1.grad, W2.grad, bits

    for _ in range(2):
        once()
    torch.cuda.synchronize()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        g1, g2, bits = once()
    torch.cuda.synchronize()
    gpu_ms = (time.perf_counter() - t0) * 1e3 / reps
    out = {"batch": 64, "steps": 200, "n_in": 1000, "n_hidden": 2000, "n_out": 100, "checkpoint_every": 20,
           "gpu_ms_per_batch": gpu_ms}
    if not args.no_cpu:
        from oracle import adjoint_ref
        T, B, H = 200, 64, 2000
        b = bits.cpu().numpy().view(np.uint64)
        u = np.unpackbits(b.view(np.uint8).reshape(T, -1), axis=1, bitorder="little")[:, :B * H]
        spikes = torch.tensor(u.reshape(T, B, H).astype(bool))
        d = {k: torch.tensor(P[k], dtype=torch.float64) for k in ("W1", "W2", "X", "target", "params")}
        W1 = d["W1"].clone().requires_grad_(True)
        W2 = d["W2"].clone().requires_grad_(True)
        t0 = time.perf_counter()
        loss, _ = adjoint_ref.forward(W1, W2, d["X"], d["target"], d["params"], P["alpha"], spikes=spikes)
        loss.backward()
        cpu_ms = (time.perf_counter() - t0) * 1e3
        rel1 = float((g1.cpu().double() - W1.grad).norm() / W1.grad.norm())
        rel2 = float((g2.cpu().double() - W2.grad).norm() / W2.grad.norm())
        out.update({"cpu_fp64_autograd_ms_per_batch": cpu_ms, "cpu_threads": torch.get_num_threads(),
                    "rel_err_dW1": rel1, "rel_err_dW2": rel2})
    log(f"[bench] adjoint: {out}")
    return out


def run_batched(args, ctx, log):
    """Batched Izhikevich instances (log-uniform sizes 500..20000, own seeds and
    noise amplitudes) in one launch vs the same instances run one after
    another as plain networks on the same GPU.  Rasters must match per instance."""
    import numpy as np
    from szsnn_amd.batch import BatchSimulation, batched_izhikevich
    from szsnn_amd.simulation import Simulation, raster_to_bool

    t0 = time.time()
    nets = batched_izhikevich(n_instances=args.batch_instances)
    n_tot = sum(n.n_neurons for n in nets)
    s_tot = sum(n.n_syn for n in nets)
    log(f"[bench] batched: {len(nets)} instances, {n_tot} neurons, {s_tot} synapses built in {time.time() - t0:.1f}s")
    T = args.batch_steps
    bs = BatchSimulation(ctx, nets)
    tiling = bs.autotune(100)
    log(f"[bench] batched tiling: {tiling}")
    bs.set_recording(0, T)
    ctx.synchronize()
    c0 = bs.counters()
    t1 = time.perf_counter()
    bs.run(T)
    ctx.synchronize()
    wall_b = time.perf_counter() - t1
    ev_b = bs.counters()["events"] - c0["events"]
    log(f"[bench] batched run: {wall_b * 1e3:.1f} ms for {T} steps, {ev_b / wall_b:.3e} events/s")
    per_inst = bs.instance_rasters()
    bs.close()
    # each instance run alone, once in graph mode and once as a persistent
    # cooperative launch; the comparison uses the faster of the two per instance
    wall = {"graph": 0.0, "persistent": 0.0, "best": 0.0}
    ev_p = 0
    same = True
    for b, net in enumerate(nets):
        best = None
        for mode in ("graph", "persistent"):
            s = Simulation(ctx, net)
            if mode == "persistent":
                s.set_mode(1, 256, 0)
            s.set_recording(0, T)
            ctx.synchronize()
            t1 = time.perf_counter()
            s.run(T)
            ctx.synchronize()
            w = time.perf_counter() - t1
            wall[mode] += w
            best = w if best is None else min(best, w)
            if mode == "graph":
                ev_p += s.counters()["events"]
            same = same and np.array_equal(raster_to_bool(s.raster(), net.n_neurons), per_inst[b])
            s.close()
        wall["best"] += best
        if b % 32 == 31:
            log(f"[bench] plain runs: {b + 1}/{len(nets)}, {wall} s so far")
    wall_p = wall["best"]
    out = {"instances": len(nets), "neurons": n_tot, "synapses": s_tot, "steps": T, "dt_ms": 1.0,
           "batched_events_per_s": ev_b / wall_b, "plain_sequential_events_per_s": ev_p / wall_p,
           "speedup": (ev_b / wall_b) / (ev_p / wall_p), "batched_wall_s": wall_b,
           "plain_sequential_wall_s": wall_p, "plain_graph_wall_s": wall["graph"],
           "plain_persistent_wall_s": wall["persistent"], "events": ev_b,
           "rasters_identical": bool(same and ev_b == ev_p), "tiling": tiling}
    log(f"[bench] batched: {out}")
    return out


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "szsnn_amd runs on one GPU; --gpus above 1 is not supported (sharding is design only)",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch
    from szsnn_amd.simulation import Context, close_all

    try:
        if args.batched_pass:
            return batched_pass(args)
        if args.adjoint_pass:
            args.no_cpu = True
            out = run_adjoint(args, log)
            print(json.dumps({"adjoint_pass": True, **out}), flush=True)
            return
        ctx = Context(0)
        net, sim = upload_headline(args, ctx)
        if args.profile_pass:
            return profile_pass(args, torch, ctx, net, sim)
        if args.burst_pass:
            return burst_pass(args, ctx, sim)
        if args.natural_pass:
            return natural_pass(args, ctx, sim)
        return headline(args, torch, ctx, net, sim)
    finally:
        # everything on the device is released here, in order, before the
        # interpreter (and a profiler attached to the HIP runtime) tears down
        close_all()
        torch.cuda.synchronize()


def upload_headline(args, ctx):
    from szsnn_amd.configs import if_cond_exp_100k
    from szsnn_amd.simulation import Simulation
    t0 = time.time()
    if args.small:
        net = if_cond_exp_100k(n_exc=8000, n_inh=2000, k_exc=400, k_inh=100)
    else:
        net = if_cond_exp_100k()
    log(f"[bench] network built in {time.time() - t0:.1f}s: {net.n_neurons} neurons, {net.n_syn} synapses")
    t0 = time.time()
    sim = Simulation(ctx, net)
    log(f"[bench] uploaded in {time.time() - t0:.1f}s to {ctx.name}; device bytes {sim.info.device_bytes}")
    return net, sim


def batched_pass(args):
    from szsnn_amd.batch import BatchSimulation, batched_izhikevich
    from szsnn_amd.simulation import Context
    ctx = Context(0)
    nets = batched_izhikevich(n_instances=args.batch_instances)
    bs = BatchSimulation(ctx, nets)
    if not args.no_autotune:
        log(f"[bench] batched tiling: {bs.autotune(100)}")
    def run(n):
        if args.batch_single:
            for _ in range(n):
                bs.run(1)
        else:
            bs.run(n)
    run(args.warmup * args.chunk)
    ctx.synchronize()
    c0 = bs.counters()
    t1 = time.perf_counter()
    run(args.batch_steps)
    ctx.synchronize()
    wall = time.perf_counter() - t1
    ev = bs.counters()["events"] - c0["events"]
    print(json.dumps({"batched_pass": True, "instances": len(nets), "neurons": sum(n.n_neurons for n in nets),
                      "steps": args.batch_steps, "us_per_step": wall * 1e6 / args.batch_steps,
                      "events_per_step": ev / args.batch_steps, "tiling": bs.tiling_ex()}), flush=True)


def profile_pass(args, torch, ctx, net, sim):
    """Exactly --volley-steps gather_kernel launches at the volley load
    (after one persistent launch has filled every delay slot), each after a
    cache-evicting sweep with --cold; a 1 GiB device copy first calibrates
    the FETCH_SIZE / WRITE_SIZE counters."""
    a = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    del a, b
    flush = CacheFlush(torch, args.flush_mib, ctx.stream) if args.cold else None
    sim.set_tiling(gather_block=args.gather_block, gather_grid=args.gather_grid)
    sim.set_mode(1 if args.fill == "persistent" else 0)
    sim.set_volley(net.n_neurons)
    sim.run(net.max_delay + 2)
    sim.set_mode(0)
    ctx.synchronize()
    c0 = sim.counters()
    if flush:
        g_ms, u_ms = time_cold(sim, flush, args.volley_steps)
        flush.close()
    else:
        g_ms, u_ms = sim.time_kernels(args.volley_steps)
    c1 = sim.counters()
    n = args.volley_steps
    ev = (c1["events"] - c0["events"]) / n
    rw = (c1["rows"] - c0["rows"]) / n
    per = gather_alg_bytes(ev, rw)
    print(json.dumps({"profile_pass": True, "cold": bool(args.cold), "gather_block": args.gather_block,
                      "gather_grid": args.gather_grid, "pull_slices": int(sim.info.pull_slices),
                      "pull_groups": int(sim.info.pull_groups), "launches": n, "gather_us": g_ms * 1e3 / n,
                      "update_us": u_ms * 1e3 / n, "events_per_launch": ev, "rows_per_launch": rw,
                      "alg_bytes_per_launch": per, "alg_gbs": per / (g_ms * 1e-3 / n) / 1e9}), flush=True)


def burst_pass(args, ctx, sim):
    sim.set_mode(0)
    heavy = sim.heavy_push()
    rows = []
    for S in (int(x) for x in args.burst_ids.split(",")):
        for name, policy, threshold in (("push_heavy", 2.0, heavy), ("push_block", 2.0, 0), ("pull", 0.0, heavy)):
            sim.set_heavy_push(threshold)
            sim.set_gather_policy(policy)
            sim.set_volley(S)
            sim.run(args.burst_steps // 10)
            ctx.synchronize()
            c0 = sim.counters()
            t1 = time.perf_counter()
            sim.run(args.burst_steps)
            ctx.synchronize()
            wall = time.perf_counter() - t1
            c1 = sim.counters()
            g_ms, u_ms = sim.time_kernels(args.volley_steps)
            rows.append({"ids": S, "mode": name, "us_per_step": wall * 1e6 / args.burst_steps,
                         "events_per_step": (c1["events"] - c0["events"]) / args.burst_steps,
                         "gather_us": g_ms * 1e3 / args.volley_steps, "update_us": u_ms * 1e3 / args.volley_steps})
            log(f"[bench] burst {rows[-1]}")
        sim.set_volley(0)
    print(json.dumps({"burst_pass": True, "heavy_threshold": heavy, "steps": args.burst_steps, "runs": rows}),
          flush=True)


def natural_pass(args, ctx, sim):
    sim.set_mode(1, 512, 256)
    sim.run(args.settle_steps + args.warmup * args.chunk)
    ctx.synchronize()
    n = args.steps * args.chunk
    c0 = sim.counters()
    t1 = time.perf_counter()
    sim.run(n)  # one persistent launch
    ctx.synchronize()
    wall = time.perf_counter() - t1
    c1 = sim.counters()
    g_ms, u_ms = sim.time_kernels(args.volley_steps)
    print(json.dumps({"natural_pass": True, "settle_steps": args.settle_steps, "steps": n,
                      "us_per_step": wall * 1e6 / n,
                      "events_per_step": (c1["events"] - c0["events"]) / n,
                      "spikes_per_step": (c1["spikes"] - c0["spikes"]) / n,
                      "single_step_gather_us": g_ms * 1e3 / args.volley_steps,
                      "single_step_update_us": u_ms * 1e3 / args.volley_steps}), flush=True)


def timed_run(sim, ctx, n, chunk=None):
    """(wall seconds, counter deltas) of sim.run(n) (or n calls of
    sim.run(chunk)), bracketed by synchronizations."""
    import torch
    torch.cuda.synchronize()
    ctx.synchronize()
    c0 = sim.counters()
    t1 = time.perf_counter()
    if chunk is None:
        sim.run(n)
    else:
        for _ in range(n):
            sim.run(chunk)
    ctx.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t1
    c1 = sim.counters()
    return wall, {k: c1[k] - c0[k] for k in ("events", "spikes", "rows", "drive_events")}


def headline(args, torch, ctx, net, sim):
    dt_ms = net.meta["dt_ms"]
    if not args.no_autotune:
        log(f"[bench] autotuned for the natural dynamics from the initial state: {sim.autotune(200)}")
    # the whole configured run (1000 ms) from the initial state: the settle
    # phase of the timed window, timed itself as the full-run figure
    wall_f, cf = timed_run(sim, ctx, args.settle_steps)
    full = {"steps": args.settle_steps, "simulated_ms": args.settle_steps * dt_ms, "wall_s": wall_f,
            "events": cf["events"], "spikes": cf["spikes"],
            "events_per_s": cf["events"] / wall_f if wall_f > 0 else None,
            "realtime_factor": args.settle_steps * dt_ms / (wall_f * 1e3) if wall_f > 0 else None,
            "firing_rate_hz": cf["spikes"] / net.n_neurons / (args.settle_steps * dt_ms * 1e-3)
            if args.settle_steps else None}
    log(f"[bench] full run: {full}")
    if not args.no_autotune:
        # tuned again on the settled dynamics (the state is restored)
        log(f"[bench] autotuned for the settled dynamics: {sim.autotune(500)}")
    sim.run(args.warmup * args.chunk)
    wall, cw = timed_run(sim, ctx, args.steps, args.chunk)
    events, spikes = cw["events"], cw["spikes"]
    n_sim = args.steps * args.chunk
    w0 = args.settle_steps + args.warmup * args.chunk
    window = f"simulation steps {w0}..{w0 + n_sim}"
    if events <= 0:
        raise SystemExit(f"bench: the timed window ({window}) delivered no synaptic event")
    ms_per_step = wall * 1e3 / args.steps
    ev_per_s = events / wall
    rtf = n_sim * dt_ms / (wall * 1e3)
    rate_hz = spikes / net.n_neurons / (n_sim * dt_ms * 1e-3)
    log(f"[bench] natural: {ms_per_step:.3f} ms per {args.chunk}-step chunk ({wall * 1e6 / n_sim:.2f} us per "
        f"simulation step), {ev_per_s:.3e} events/s, RTF {rtf:.2f}, {rate_hz:.2f} Hz, "
        f"{events / n_sim:.0f} events per simulation step")

    # per-kernel time on the natural dynamics (HIP events, same stream); the
    # update kernel also pushes the step's spikes
    kn = 500
    cg0 = sim.counters()
    g_ms, u_ms = sim.time_kernels(kn)
    cg1 = sim.counters()
    info = sim.info
    # per neuron: params, vars read+written, its accumulator slot, key,
    # synapse range and delay bits; per pushed synapse: its 8-byte entry
    nat_u_bytes = ((4 * info.n_params + 8 * info.n_vars + 16 + 24) * net.n_neurons * kn
                   + 8 * (cg1["events"] - cg0["events"]))
    natural = {"gather_us": g_ms * 1e3 / kn, "update_us": u_ms * 1e3 / kn,
               "update_gbs": nat_u_bytes / (u_ms * 1e-3) / 1e9 if u_ms > 0 else None}
    log(f"[bench] natural kernels: {natural}")
    natural_tiling = (sim.tiling(), sim.mode())

    default_policy = sim.gather_policy()["dense_fraction"]
    groups = int(sim.info.pull_groups)
    pull_entry_bytes = int(sim.info.pull_entry_bytes)
    flush = CacheFlush(torch, args.flush_mib, ctx.stream)

    def volley(fraction, cold=False, ids=None):
        sim.set_gather_policy(fraction)
        sim.set_volley(net.n_neurons if ids is None else ids)
        sim.run(net.max_delay + 2)  # fill all delay slots
        if not args.no_autotune:
            log(f"[bench] autotuned for the volley (policy {fraction}): {sim.autotune(10)}")
        sim.set_mode(0)  # time the kernels one launch at a time
        cv0 = sim.counters()
        if cold:
            gv_ms, uv_ms = time_cold(sim, flush, args.volley_steps)
        else:
            gv_ms, uv_ms = sim.time_kernels(args.volley_steps)
        cv1 = sim.counters()
        sim.set_volley(0)
        ev = (cv1["events"] - cv0["events"]) / args.volley_steps
        rw = (cv1["rows"] - cv0["rows"]) / args.volley_steps
        per = gather_alg_bytes(ev, rw)
        g_s = gv_ms * 1e-3 / args.volley_steps
        return {"gather_us": g_s * 1e6, "update_us": uv_ms * 1e3 / args.volley_steps, "events_per_launch": ev,
                "rows_per_launch": rw, "alg_bytes_per_launch": per, "alg_gbs": per / g_s / 1e9,
                "tiling": sim.tiling()}
    # push only: 1e8 synapses per step added with atomics (every update block
    # holds 512k synapse entries, so it posts them as heavy pushes and the
    # next gather launch adds them; the step is both kernels)
    v_push = volley(2.0)
    v_push["push_events_per_s"] = v_push["events_per_launch"] / ((v_push["update_us"] + v_push["gather_us"]) * 1e-6)
    log(f"[bench] volley push-only: {v_push}")
    # default policy: every step is deferred and pulled; warm (launches back
    # to back: the 244 MB pull layout largely stays in the Infinity Cache)
    # and cold (a cache-evicting sweep before every launch)
    v_warm = volley(default_policy)
    v_warm["pull_groups"] = groups
    v_warm["pull_slices"] = int(sim.info.pull_slices)
    log(f"[bench] volley pull, warm (policy {default_policy}): {v_warm}")
    v_cold = volley(default_policy, cold=True)
    log(f"[bench] volley pull, cold: {v_cold}")
    # the volley benchmark at 1 %, 10 % and 100 % of the neurons forced every
    # step (ids 0 .. k - 1; the network's own spikes come on top), warm, one
    # launch of each kernel at a time: under the default policy (a step is
    # pulled when its predicted spikes, the previous step's, reach the
    # policy's fraction of the ids with synapses, else pushed) and with every
    # step pulled (policy 0)
    fractions = []
    for f in (0.01, 0.1, 1.0):
        k = int(round(f * net.n_neurons))
        row = {"fraction": f, "forced_ids": k}
        for name, pol in (("default_policy", default_policy), ("pull_every_step", 0.0)):
            v = volley(pol, ids=k)
            step_us = v["gather_us"] + v["update_us"]
            row[name] = {"events_per_step": v["events_per_launch"], "rows_per_step": v["rows_per_launch"],
                         "gather_us": v["gather_us"], "update_us": v["update_us"], "us_per_step": step_us,
                         "events_per_s": v["events_per_launch"] / (step_us * 1e-6),
                         "alg_gbs_per_step": v["alg_bytes_per_launch"] / (step_us * 1e-6) / 1e9}
        fractions.append(row)
        log(f"[bench] volley {f:.0%}: {row}")
    sim.set_gather_policy(default_policy)
    flush.close()
    sim.close()

    copy_gbs = copy_bandwidth_gbs(torch)
    log(f"[bench] device copy: {copy_gbs:.0f} GB/s (read + write)")
    t_warm = read_json(args.traffic_file)
    t_cold = read_json(args.cold_traffic_file)
    batched = None if args.no_batched else run_batched(args, ctx, log)
    adjoint = None if args.no_adjoint else run_adjoint(args, log)

    cpu = None
    if not args.no_cpu:
        # the FP32 CPU restatement of the same network: the same settle run
        # (timed: the CPU's full-run rate), warm-up, and the GPU's timed
        # window, on every core this process may use; then a one-thread
        # sample of the steps after it
        from oracle.pyoracle import Oracle
        threads = len(os.sched_getaffinity(0))
        t0 = time.time()
        o = Oracle(net, 32)
        log(f"[bench] CPU oracle built in {time.time() - t0:.1f}s")
        secs_f = o.time_steps(args.settle_steps, threads=threads)
        ef = o.events
        o.time_steps(args.warmup * args.chunk, threads=threads)
        e0 = o.events
        secs_w = o.time_steps(n_sim, threads=threads)
        e1 = o.events
        secs_1 = o.time_steps(args.cpu_steps)
        e2 = o.events
        cpu = {"value": (e1 - e0) / secs_w, "unit": "events/s", "cores": threads, "kind": "port",
               "sample": f"the GPU's timed window ({window}, {n_sim * dt_ms:.0f} ms simulated) of the same network "
                         f"after the same settle run and warm-up, FP32 CPU restatement on {threads} OpenMP threads: "
                         f"{secs_w:.2f} s wall, RTF {n_sim * dt_ms / (secs_w * 1e3):.3f}",
               "events_match_gpu": int(e1 - e0) == int(events),
               "full_run_events_per_s": ef / secs_f,
               "full_run_wall_s": secs_f,
               "full_run_events_match_gpu": int(ef) == int(full["events"]),
               "value_1thread": (e2 - e1) / secs_1,
               "sample_1thread": f"the next {args.cpu_steps} steps on one thread: {secs_1:.2f} s wall, RTF "
                                 f"{args.cpu_steps * dt_ms / (secs_1 * 1e3):.3f}",
               "host": platform.processor() or platform.machine(), "host_cpus": os.cpu_count()}
        log(f"[bench] CPU: {cpu}")
        del o

    g_cold_s = v_cold["gather_us"] * 1e-6
    g_warm_s = v_warm["gather_us"] * 1e-6
    traffic = t_cold["hbm_bytes_per_launch"] if t_cold else None
    achieved = traffic / g_cold_s / 1e9 if traffic else None
    roofline = {
        "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": traffic,
        "kernel": "gather_kernel pull, 2-byte count-format entries (volley: all 100k neurons forced each step), "
                  "cold: a 2 GiB read sweep before each launch",
        "avg_launch_us": g_cold_s * 1e6,
        "timing": "HIP events recorded by each launch's dispatch on the library's stream (hipExtLaunchKernelGGL)",
        "traffic_detail": t_cold,
        "alg_bytes_per_launch": v_cold["alg_bytes_per_launch"], "alg_gbs": v_cold["alg_gbs"],
        "alg_unit": "8 B per delivered synapse + 8 B per delivered (source, delay) row",
        "pull_entry_bytes": pull_entry_bytes, "copy_gbs": copy_gbs,
        "warm": {"avg_launch_us": g_warm_s * 1e6,
                 "traffic": t_warm["hbm_bytes_per_launch"] if t_warm else None,
                 "achieved": t_warm["hbm_bytes_per_launch"] / g_warm_s / 1e9 if t_warm else None,
                 "frac": t_warm["hbm_bytes_per_launch"] / g_warm_s / 1e9 / HBM_PEAK_GBS if t_warm else None,
                 "alg_gbs": v_warm["alg_gbs"]},
    }
    out = {
        "metric": "synaptic events/s (plain IF_cond_exp 100k net); realtime factor; gather % HBM peak",
        "value": ev_per_s,
        "unit": "events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (int64 fixed-point accumulation)",
        "data": "synthetic (seeded random network and drive)",
        "config": {"workload": "if_cond_exp_100k" if not args.small else "if_cond_exp_10k_small",
                   "neurons": net.n_neurons, "synapses": net.n_syn, "dt_ms": dt_ms,
                   "max_delay_steps": net.max_delay, "drive": "Binomial(100, 5 Hz * dt) per neuron-step",
                   "settle_steps": args.settle_steps, "simulation_steps_per_bench_step": args.chunk},
        "realtime_factor": rtf,
        "firing_rate_hz": rate_hz,
        "events_per_step": events / args.steps,
        "events_per_simulation_step": events / n_sim,
        "us_per_simulation_step": wall * 1e6 / n_sim,
        "full_run": full,
        "natural_kernels": natural,
        "natural_tiling": {**natural_tiling[0], **natural_tiling[1]},
        "volley_push_only": v_push,
        "volley_pull_warm": v_warm,
        "volley_pull_cold": v_cold,
        "volley_fractions": fractions,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "batched": batched,
        "adjoint": adjoint,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
