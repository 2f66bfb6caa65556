"""pyszsnn host interface: Context and Simulation over the C ABI.

Mirrors the reference's C++ simulation API (szsnn1.1/src/Simulation.hpp:39-131):

    ctx = Context()                      # Context::setup (Context.cpp:136-224)
    sim = Simulation(ctx, network)       # Simulation(const Context&, NetworkGraph&)
    sim.step() / sim.step(fInput)        # Simulation.ipp:471-496
    ids = sim.stepAndFired(fInput)       # Simulation.ipp:500-530, ascending neuron ids
    sim.getTime()                        # wall-clock ms spent stepping (SzsnnTimer)

plus the batched-run calls of the HIP library (run, recording, counters,
volley input, tiling).  Errors raise SzsnnError, the reference's
SzsnnException contract (utils.h:144-155).
"""
import atexit
import ctypes
import time
import weakref

import numpy as np

from . import _lib
from ._lib import SzsnnError, check, ptr


# Live contexts and networks.  At interpreter exit they are closed in
# dependency order (networks, then contexts), before the HIP runtime's own
# exit handlers run: left to garbage collection at shutdown, a network could
# free device memory and graphs after the runtime (or a profiler attached to
# it) has already been torn down.
_live_nets = weakref.WeakSet()
_live_ctxs = weakref.WeakSet()


@atexit.register
def close_all():
    """Close every live Simulation, then every live Context."""
    for s in list(_live_nets):
        try:
            s.close()
        except Exception:
            pass
    for c in list(_live_ctxs):
        try:
            c.close()
        except Exception:
            pass


class Context:
    """One device and one HIP stream (replaces Context, Context.hpp:77-303)."""

    def __init__(self, device=0):
        self.lib = _lib.load()
        h = ctypes.c_void_p()
        check(self.lib.szsnn_create(int(device), ctypes.byref(h)))
        self.handle = h
        self.device = device
        _live_ctxs.add(self)

    @property
    def name(self):
        buf = ctypes.create_string_buffer(256)
        check(self.lib.szsnn_device_name(self.handle, buf, 256), self.handle)
        return buf.value.decode()

    def synchronize(self):
        check(self.lib.szsnn_synchronize(self.handle), self.handle)

    @property
    def stream(self):
        """The context's hipStream_t, as an integer (for torch.cuda.ExternalStream)."""
        return int(self.lib.szsnn_stream(self.handle) or 0)

    def close(self):
        """Destroy the stream; networks still open on this context are closed first."""
        if self.handle:
            for s in list(_live_nets):
                if s.ctx is self:
                    s.close()
            self.lib.szsnn_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _desc(net):
    """Fill szsnn_net_desc from a Network; returns (desc, keepalive)."""
    keep = []

    def c(a, dtype):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=dtype)
        keep.append(a)
        return a

    d = _lib.NetDesc()
    d.model = net.model
    d.n_neurons = net.n_neurons
    d.n_sources = net.n_sources
    d.source_model = net.source_model
    d.params = ptr(c(net.params, np.float32), ctypes.c_float)
    d.vars = ptr(c(net.vars, np.float32), ctypes.c_float)
    d.src_rate = ptr(c(net.src_rate, np.float32), ctypes.c_float)
    d.src_last = ptr(c(net.src_last, np.float32), ctypes.c_float)
    d.src_thresh = ptr(c(net.src_thresh, np.uint64), ctypes.c_uint64)
    d.n_syn = net.n_syn
    d.syn_src = ptr(c(net.syn_src, np.uint32), ctypes.c_uint32)
    d.syn_tgt = ptr(c(net.syn_tgt, np.uint32), ctypes.c_uint32)
    d.syn_w = ptr(c(net.syn_w, np.float32), ctypes.c_float)
    d.syn_delay = ptr(c(net.syn_delay, np.uint16), ctypes.c_uint16)
    d.max_delay = net.max_delay
    d.frac_bits = net.frac_bits
    d.seed = net.seed & 0xFFFFFFFFFFFFFFFF
    d.noise_amp = ptr(c(net.noise_amp, np.float32), ctypes.c_float)
    cdf = c(net.drive_cdf, np.uint64)
    d.drive_k = 0 if cdf is None else int(cdf.shape[0])
    d.drive_cdf = ptr(cdf, ctypes.c_uint64)
    d.drive_w = net.drive_w
    d.queue_capacity = net.queue_capacity
    d.pull_layout = getattr(net, "pull_layout", 0)
    return d, keep


class Simulation:
    """A network uploaded to one device, stepped by the HIP library."""

    def __init__(self, context, network):
        if isinstance(context, int):
            context = Context(context)
        self.ctx = context
        self.lib = context.lib
        self.network = network
        d, keep = _desc(network)
        h = ctypes.c_void_p()
        check(self.lib.szsnn_net_upload(context.handle, ctypes.byref(d), ctypes.byref(h)), context.handle)
        del keep
        self.handle = h
        info = _lib.NetInfo()
        check(self.lib.szsnn_net_get_info(h, ctypes.byref(info)), context.handle)
        self.info = info
        self.n_neurons = info.n_neurons
        self.words = int(info.raster_words)
        self._wall_ms = 0.0
        self._rec = None
        self._bitmap = np.zeros(self.words, np.uint64)
        _live_nets.add(self)

    # -- reference API -------------------------------------------------
    def step(self, fInput=None):
        self._step(fInput, None)

    def stepAndFired(self, fInput=None):
        """Spiking neuron ids of this step, ascending (SpikeBuffer::getOutput
        with SZSNN_NO_INPUT_TO_OUTPUT=1: neurons only)."""
        self._step(fInput, self._bitmap)
        ids = bitmap_to_ids(self._bitmap)
        return ids[ids < self.n_neurons]

    def stepAndFiredAll(self, fInput=None):
        """Like stepAndFired but including input-source ids (N..N+M-1)."""
        self._step(fInput, self._bitmap)
        return bitmap_to_ids(self._bitmap)

    def getTime(self):
        return self._wall_ms

    # single-kernel harnesses (Simulation.ipp:142-210): one kernel of the
    # current step, launched alone and waited for; the cycle does not advance
    def testUpdate(self):
        """The update kernel of the current step (with the scatter of its
        spikes); returns its time in ms."""
        return self._test_kernel(_lib.KERNEL_UPDATE)

    def testScatter(self):
        """The scatter runs inside the update kernel here: the same launch as
        testUpdate."""
        return self._test_kernel(_lib.KERNEL_UPDATE)

    def testGather(self):
        """The gather kernel of the current step; returns its time in ms."""
        return self._test_kernel(_lib.KERNEL_GATHER)

    def _test_kernel(self, which):
        ms = ctypes.c_double()
        check(self.lib.szsnn_test_kernel(self.handle, int(which), ctypes.byref(ms)), self.ctx.handle)
        return ms.value

    # -- batched API ----------------------------------------------------
    def run(self, n_steps, sync=True):
        t0 = time.perf_counter()
        fn = self.lib.szsnn_run if sync else self.lib.szsnn_run_async
        check(fn(self.handle, int(n_steps)), self.ctx.handle)
        self._wall_ms += (time.perf_counter() - t0) * 1e3

    def set_recording(self, first, n_steps):
        check(self.lib.szsnn_set_recording(self.handle, int(first), int(n_steps)), self.ctx.handle)
        self._rec = (int(first), int(n_steps))

    def raster(self):
        """Recorded bitmaps as uint64 [rec_steps, words]."""
        if self._rec is None:
            raise SzsnnError(_lib.SZSNN_ERR_STATE, "recording is off")
        out = np.zeros((self._rec[1], self.words), np.uint64)
        check(self.lib.szsnn_read_raster(self.handle, ptr(out, ctypes.c_uint64)), self.ctx.handle)
        return out

    def set_volley(self, n_prefix):
        check(self.lib.szsnn_set_volley(self.handle, int(n_prefix)), self.ctx.handle)

    def get_vars(self):
        nv = self.info.n_vars
        out = np.zeros((nv, self.n_neurons), np.float32)
        check(self.lib.szsnn_get_vars(self.handle, ptr(out, ctypes.c_float)), self.ctx.handle)
        return out

    def set_vars(self, vars_):
        v = np.ascontiguousarray(vars_, np.float32)
        assert v.shape == (self.info.n_vars, self.n_neurons)
        check(self.lib.szsnn_set_vars(self.handle, ptr(v, ctypes.c_float)), self.ctx.handle)

    def counters(self):
        c = _lib.Counters()
        check(self.lib.szsnn_get_counters(self.handle, ctypes.byref(c)), self.ctx.handle)
        return {"cycle": c.cycle, "events": c.events, "spikes": c.spikes,
                "drive_events": c.drive_events, "overflow": c.overflow, "rows": c.rows,
                "deferred": c.deferred, "late_deferred": c.late_deferred}

    def reset_counters(self):
        check(self.lib.szsnn_reset_counters(self.handle), self.ctx.handle)

    def time_kernels(self, n_steps):
        g = ctypes.c_double()
        u = ctypes.c_double()
        check(self.lib.szsnn_time_kernels(self.handle, int(n_steps), ctypes.byref(g), ctypes.byref(u)),
              self.ctx.handle)
        return g.value, u.value

    def set_tiling(self, update_block=0, gather_block=0, gather_grid=0):
        check(self.lib.szsnn_set_tiling(self.handle, update_block, gather_block, gather_grid), self.ctx.handle)

    def tiling(self):
        a, b, c = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        check(self.lib.szsnn_get_tiling(self.handle, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)),
              self.ctx.handle)
        return {"update_block": a.value, "gather_block": b.value, "gather_grid": c.value}

    def tiling_ex(self):
        """The whole launch configuration (szsnn_get_tiling_ex) as a dict."""
        t = _lib.Tiling()
        check(self.lib.szsnn_get_tiling_ex(self.handle, ctypes.byref(t)), self.ctx.handle)
        return {name: getattr(t, name) for name, _ in _lib.Tiling._fields_}

    def set_tiling_ex(self, **changes):
        """Change fields of the launch configuration (the others are kept);
        gather_stage 0 stages as many delays as the LDS admits."""
        cur = self.tiling_ex()
        unknown = set(changes) - set(cur) | ({"gather_stage_max"} & set(changes))
        if unknown:
            raise ValueError(f"unknown or read-only tiling fields: {sorted(unknown)}")
        cur.update(changes)
        t = _lib.Tiling(**{k: v for k, v in cur.items()})
        check(self.lib.szsnn_set_tiling_ex(self.handle, ctypes.byref(t)), self.ctx.handle)

    def set_gather_policy(self, dense_fraction):
        """Defer a step's spikes to the pull gather when the spikes predicted
        for it (the previous step's less the ids forced in it, plus the ids
        forced in this one) reach dense_fraction x the ids with synapses
        (<= 0: always, > 1: never).  The default follows the pull entry
        width: 0.012 (2-byte), 0.02 (4-byte), 0.04 (8-byte)."""
        check(self.lib.szsnn_set_gather_policy(self.handle, float(dense_fraction)), self.ctx.handle)

    def gather_policy(self):
        f, h = ctypes.c_float(), ctypes.c_int32()
        check(self.lib.szsnn_get_gather_policy(self.handle, ctypes.byref(f), ctypes.byref(h)), self.ctx.handle)
        return {"dense_fraction": f.value, "has_pull": bool(h.value)}

    def set_heavy_push(self, min_entries):
        """On a pushed step, an update block whose spiking ids carry at least
        min_entries synapse entries hands them to the next gather launch,
        which adds them with its whole grid (0: every block pushes its own).
        Needs the pull layout unless min_entries is 0."""
        check(self.lib.szsnn_set_heavy_push(self.handle, int(min_entries)), self.ctx.handle)

    def heavy_push(self):
        m = ctypes.c_uint32()
        check(self.lib.szsnn_get_heavy_push(self.handle, ctypes.byref(m)), self.ctx.handle)
        return m.value

    def set_mode(self, mode, block=0, grid=0):
        check(self.lib.szsnn_set_mode(self.handle, int(mode), int(block), int(grid)), self.ctx.handle)

    def mode(self):
        m, b, g = ctypes.c_int32(), ctypes.c_uint32(), ctypes.c_uint32()
        check(self.lib.szsnn_get_mode(self.handle, ctypes.byref(m), ctypes.byref(b), ctypes.byref(g)),
              self.ctx.handle)
        return {"mode": m.value, "block": b.value, "grid": g.value}

    def autotune(self, steps=100):
        check(self.lib.szsnn_autotune(self.handle, int(steps)), self.ctx.handle)
        return self.tiling_ex()

    def close(self):
        if getattr(self, "handle", None):
            self.lib.szsnn_net_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- internals -------------------------------------------------------
    def _step(self, fInput, out):
        t0 = time.perf_counter()
        if fInput is not None and len(fInput):
            ids = np.ascontiguousarray(fInput, np.uint32)
            check(self.lib.szsnn_step(self.handle, ptr(ids, ctypes.c_uint32), ids.shape[0],
                                      ptr(out, ctypes.c_uint64)), self.ctx.handle)
        else:
            check(self.lib.szsnn_step(self.handle, None, 0, ptr(out, ctypes.c_uint64)), self.ctx.handle)
        self._wall_ms += (time.perf_counter() - t0) * 1e3


def bitmap_to_ids(bits):
    """Ascending ids of set bits in a uint64 bitmap (SpikeBuffer::bitvectorToSparse)."""
    b = np.ascontiguousarray(bits, np.uint64).view(np.uint8)
    return np.flatnonzero(np.unpackbits(b, bitorder="little")).astype(np.uint32)


def raster_to_bool(raster, n_ids):
    """[steps, words] uint64 bitmaps -> [steps, n_ids] bool."""
    r = np.ascontiguousarray(raster, np.uint64)
    bits = np.unpackbits(r.view(np.uint8).reshape(r.shape[0], -1), axis=1, bitorder="little")
    return bits[:, :n_ids].astype(bool)
